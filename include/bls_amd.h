/*
 * bls_amd.h -- C ABI of the MI355X elimination library (libblsamd.so).
 *
 * Each entry point replaces the reference function cited beside it (BlockLinearSystems.jl, file:line) and takes
 * the reference's own arrays verbatim: index arrays are Int64 holding the 1-based values the Julia structures
 * store, blocks are column-major, C_er is stored d_e x d_r (src/schur_assembly.jl:236-246).
 *
 * Memory: integer maps are copied to the device once, when a handle is created (the structure is immutable).
 * Value arrays (E_data, C_data, R, b, S, g_S, tables, ...) are DEVICE pointers; callers with host arrays stage
 * them with bls_alloc / bls_upload / bls_download. All work of a call is enqueued on `stream` (a hipStream_t;
 * NULL = the null stream). A handle may be used from one host thread at a time; distinct handles may be used
 * concurrently on distinct streams.
 *
 * Return codes: BLS_OK (0) success; BLS_FAILED (1) numerical failure, the reference's `false`, with the lowest
 * failing eliminated block (1-based) in *failed_block; negative values are errors, described by bls_last_error().
 * Argument checks the reference makes before any work (check_damping, src/utils.jl:100-107; reduce_qr!'s
 * structure and length checks, src/qr_reduction.jl:166-169) are made here on scalars and sizes; checks over
 * device arrays (finite, non-negative scaling vectors) are the caller's, made on its host copies before the call.
 */
#ifndef BLS_AMD_H
#define BLS_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLS_OK 0
#define BLS_FAILED 1
#define BLS_ERR_ARG (-1)
#define BLS_ERR_HIP (-2)
#define BLS_ERR_UNSUPPORTED (-3)
#define BLS_ERR_STATE (-4)

typedef struct bls_schur bls_schur;       /* a registered SchurStructure + its reduced-system layout */
typedef struct bls_qr bls_qr;             /* a registered SchurJacobianStructure */
typedef struct bls_jfill bls_jfill;       /* a SchurJacobianStructure's row blocks, for filling its tables */

/* Message of the last error on this host thread ("" if none). */
const char *bls_last_error(void);
/* Library version string. */
const char *bls_version(void);

/* ---- device memory helpers (for callers holding host arrays) ---------------------------------------- */
int bls_alloc(void **ptr, int64_t bytes);
int bls_free(void *ptr);
int bls_upload(void *dst_device, const void *src_host, int64_t bytes, void *stream);
int bls_download(void *dst_host, const void *src_device, int64_t bytes, void *stream);
int bls_stream_synchronize(void *stream);

/* ---- Schur path ------------------------------------------------------------------------------------- */

/* Register a SchurStructure (src/schur_structure.jl:50-68) with the retained-block-major index of its
 * SchurReductionWorkspace (src/schur_reduction.jl:51-92). r_dims / e_dims: the block sizes (collect of the
 * BlockRLEs); boffsets_r, boffsets_e, eoffsets, eliminated_ptr, entry_retained, coffsets: the structure's fields;
 * retained_ptr, retained_entry: the workspace's. Eliminated blocks of d_e <= 8 (the reference's static path,
 * src/kernels.jl:88) with retained blocks of d_r <= 16 take the register kernels; an instance holding any larger
 * block (d_e up to 4096, any d_r) is reduced whole by the run-time-size kernels, as the reference's dynamic path
 * (src/schur_reduction.jl:164-225, :578-614) treats such blocks. d_e above 4096 returns BLS_ERR_UNSUPPORTED. */
int bls_schur_create(int64_t nr, int64_t nrblocks, int64_t neblocks, int64_t ncoup, const int64_t *r_dims,
                     const int64_t *e_dims, const int64_t *boffsets_r, const int64_t *boffsets_e,
                     const int64_t *eoffsets, const int64_t *eliminated_ptr, const int64_t *entry_retained,
                     const int64_t *coffsets, const int64_t *retained_ptr, const int64_t *retained_entry,
                     bls_schur **out);
void bls_schur_destroy(bls_schur *h);

/* Attach the sparse reduced-system layout of build_schur_pattern (src/schur_structure.jl:241-323): S.colptr
 * (nr + 1), the block pattern (block_colptr nrblocks + 1, block_rowval nnz_blocks, block_row_offsets
 * nnz_blocks, col_heights nrblocks), and for reset_schur_sparse! the r_to_s_map of build_r_to_s_map
 * (src/schur_structure.jl:406-431, length nR) and diag_inds (nr). */
int bls_schur_set_sparse(bls_schur *h, int64_t nnz, const int64_t *S_colptr, int64_t nnz_blocks,
                         const int64_t *block_colptr, const int64_t *block_rowval, const int64_t *block_row_offsets,
                         const int64_t *col_heights, int64_t nR, const int64_t *r_to_s_map, const int64_t *diag_inds);
/* Use a dense nr x nr column-major S (reduce_schur_dense!). */
int bls_schur_set_dense(bls_schur *h);

/* reset_schur_sparse! (src/schur_reduction.jl:416-436): nzval = 0; nzval[r_to_s_map] = R_nzval; scaled by s_r
 * (NULL = UnitScale); nzval[diag_inds] += lambda * scaling_r * s_r^2. */
int bls_reset_schur_sparse_f64(bls_schur *h, double *nzval, const double *R_nzval, const double *scaling_r,
                               double lambda, const double *s_r, void *stream);
/* reset_schur_dense! (src/schur_reduction.jl:389-407). */
int bls_reset_schur_dense_f64(bls_schur *h, double *S, const double *R, const double *scaling_r, double lambda,
                              const double *s_r, void *stream);

/* reduce_schur_sparse! (src/schur_reduction.jl:489-507) / reduce_schur_dense! (:465-479): phase 1 caches
 * E_inv (sum d_e^2) and z (sum d_e) -- the workspace's E_inv_data and z_data --, then g_S = b_r .* s_r and the
 * phase-2 updates of S (upper triangle) and g_S, in the reference's order (per S entry, ascending eliminated
 * block), bitwise reproducible. Synchronizes `stream` to return the status. */
int bls_reduce_schur_sparse_f64(bls_schur *h, double *nzval, double *g_S, const double *b_r, const double *b_e,
                                const double *scaling_e, double lambda, const double *E_data, const double *C_data,
                                double *E_inv, double *z, const double *s_r, const double *s_e,
                                int64_t *failed_block, void *stream);
int bls_reduce_schur_dense_f64(bls_schur *h, double *S, double *g_S, const double *b_r, const double *b_e,
                               const double *scaling_e, double lambda, const double *E_data, const double *C_data,
                               double *E_inv, double *z, const double *s_r, const double *s_e,
                               int64_t *failed_block, void *stream);
/* The same without the synchronization: the lowest failing block (0 = none) is written to the device int64
 * *d_failed_block when the stream reaches it. For pipelines and graph capture. */
int bls_reduce_schur_sparse_f64_async(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                      const double *b_e, const double *scaling_e, double lambda,
                                      const double *E_data, const double *C_data, double *E_inv, double *z,
                                      const double *s_r, const double *s_e, int64_t *d_failed_block, void *stream);
int bls_reduce_schur_dense_f64_async(bls_schur *h, double *S, double *g_S, const double *b_r, const double *b_e,
                                     const double *scaling_e, double lambda, const double *E_data,
                                     const double *C_data, double *E_inv, double *z, const double *s_r,
                                     const double *s_e, int64_t *d_failed_block, void *stream);
/* reduce_schur_dense! (src/schur_reduction.jl:465-479, accumulate_dense_columns! :262-313) with phase 2 as a SYRK on
 * the matrix cores, for dense reduced systems of many landmarks (no conditioning scales): phase 1 as in
 * bls_reduce_schur_dense_f64 (E_inv, z, the lowest failing block; S and g_S untouched on failure), then with
 * E_inv_e = M M' and Y_e = M' C_e, S -= Y' Y over the (sum d_e) x nr matrix Y of all landmarks: upper 128 x 128 tiles
 * on MFMA 16x16x4 in K slices of Y's rows (nslices; 0 picks one), each slice's tile accumulated in the element type
 * of Y, the slices added into S in Float64 in slice order; g_S = b_r - sum C' z in Float64. _f64: Float64 E, C, Y
 * and FP64 MFMA. _mixed: Float64 E (the landmark factorisation stays Float64), Float32 couplings C -- the bulk of
 * the data --, Y rounded to Float32 and FP32 MFMA with Float32 accumulation per tile and slice (the FP32-MFMA SYRK
 * with Float64 cross-tile accumulation). Holds Y and
 * the partial tiles in the handle (allocated at the first call). Synchronizes. Blocks up to d_e 8 and d_r 16
 * (BLS_ERR_UNSUPPORTED otherwise). */
int bls_reduce_schur_dense_mfma_f64(bls_schur *h, double *S, double *g_S, const double *b_r, const double *b_e,
                                    const double *scaling_e, double lambda, const double *E_data, const double *C_data,
                                    double *E_inv, double *z, int64_t nslices, int64_t *failed_block, void *stream);
int bls_reduce_schur_dense_mfma_mixed(bls_schur *h, double *S, double *g_S, const double *b_r, const double *b_e,
                                      const double *scaling_e, double lambda, const double *E_data, const float *C_data,
                                      double *E_inv, double *z, int64_t nslices, int64_t *failed_block, void *stream);
/* factor_all_eliminated! (src/schur_reduction.jl:439-455): phase 1 alone -- E_inv and z of every eliminated
 * block in Float64, the lowest failing block (1-based) in *failed_block and BLS_FAILED if one fails; synchronizes. */
int bls_schur_factor_f64(bls_schur *h, const double *E_data, const double *b_e, const double *scaling_e, double lambda,
                         double *E_inv, double *z, const double *s_e, int64_t *failed_block, void *stream);
/* The retained rows of the full system's residual after a back-substitution: out = b_r .* s_r - sum_c scaled(C_c)'
 * x_e[e(c)] per retained block, its couplings in ascending eliminated block (fixed-order sums, no atomics). With
 * x_e = z it is g_S of the reduction; with x_e = backsubstitute!(x_r) it is g_S - (S_red - R_damped) x_r. No
 * reference counterpart (the Float64 iterative refinement of the mixed reduction); C in Float64 or Float32. Blocks
 * up to d_e 8 and d_r 16 (BLS_ERR_UNSUPPORTED otherwise). */
int bls_schur_retained_residual_f64(bls_schur *h, double *out, const double *b_r, const double *x_e,
                                    const double *C_data, const double *s_r, const double *s_e, void *stream);
int bls_schur_retained_residual_f32(bls_schur *h, double *out, const double *b_r, const double *x_e,
                                    const float *C_data, const double *s_r, const double *s_e, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops): bls_reduce_schur_sparse_f64_async without conditioning scales (s_r = s_e =
 * UnitScale), whose scratch -- the failure flag and the row sums of split items -- is a static array of the
 * library (16 MiB per GPU, held until the process ends) instead of the handle's buffers. Makes no allocation.
 * Returns BLS_ERR_ARG when the handle's work list needs more scratch than that (longer parts, set with
 * bls_schur_set_tuning, need less). Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams, or from different host threads) race; serialize them. */
int bls_reduce_schur_sparse_f64_noalloc(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                        const double *b_e, const double *scaling_e, double lambda,
                                        const double *E_data, const double *C_data, double *E_inv, double *z,
                                        int64_t *d_failed_block, void *stream);

/* backsubstitute! (src/schur_reduction.jl:519-576): x_e = z_e - E_inv sum_r C_er x_r, from the E_inv and z of
 * the last reduction. */
int bls_backsubstitute_schur_f64(bls_schur *h, double *x_e, const double *x_r, const double *C_data,
                                 const double *E_inv, const double *z, const double *s_r, const double *s_e,
                                 void *stream);

/* copyto!(dest::SparseMatrixCSC, src, scatter_map) (src/SparseBlockLinearSystem.jl:356-362): the gather
 * dest[i] = src[scatter_map[i]] (1-based map, host array, uploaded at each call unless a device copy is given
 * by bls_scatter_map_create). */
int bls_copyto_csc_f64(int64_t n, double *dest, const double *src, const int64_t *scatter_map_device_i64,
                       void *stream);
/* ---- Float32 and mixed variants of the Schur path.
 * _f32: SchurBlockHessian{Float32} (src/schur_hessian.jl:31-77 with T = Float32; the element-type rules of
 * src/schur_linear_systems.jl:13-21): E, C, R, S and E_inv in Float32; b, z, g_S, the damping and the scales in
 * Float64. The landmark blocks are inverted in Float32 (inv_sym_block with T = Float32, src/kernels.jl:14-83), so a
 * block is rejected exactly when the reference's Float32 path rejects it; the phase-2 products are formed and
 * accumulated in Float64 registers and S is rounded to Float32 once.
 * _mixed (no reference counterpart; for callers holding Float32 blocks who want the Float64 reduced system): E, C
 * and R in Float32; the landmark factorisation, E_inv, S and g_S in Float64. S is then the Schur complement of the
 * stored Float32 blocks to Float64 rounding. Arguments are those of the _f64 entry points. */
int bls_reset_schur_sparse_f32(bls_schur *h, float *nzval, const float *R_nzval, const double *scaling_r,
                               double lambda, const double *s_r, void *stream);
int bls_reset_schur_dense_f32(bls_schur *h, float *S, const float *R, const double *scaling_r, double lambda,
                              const double *s_r, void *stream);
int bls_reduce_schur_sparse_f32(bls_schur *h, float *nzval, double *g_S, const double *b_r, const double *b_e,
                                const double *scaling_e, double lambda, const float *E_data, const float *C_data,
                                float *E_inv, double *z, const double *s_r, const double *s_e,
                                int64_t *failed_block, void *stream);
int bls_reduce_schur_dense_f32(bls_schur *h, float *S, double *g_S, const double *b_r, const double *b_e,
                               const double *scaling_e, double lambda, const float *E_data, const float *C_data,
                               float *E_inv, double *z, const double *s_r, const double *s_e,
                               int64_t *failed_block, void *stream);
int bls_reduce_schur_sparse_f32_async(bls_schur *h, float *nzval, double *g_S, const double *b_r,
                                      const double *b_e, const double *scaling_e, double lambda,
                                      const float *E_data, const float *C_data, float *E_inv, double *z,
                                      const double *s_r, const double *s_e, int64_t *d_failed_block, void *stream);
int bls_backsubstitute_schur_f32(bls_schur *h, double *x_e, const double *x_r, const float *C_data,
                                 const float *E_inv, const double *z, const double *s_r, const double *s_e,
                                 void *stream);
int bls_copyto_csc_f32(int64_t n, float *dest, const float *src, const int64_t *scatter_map_device_i64,
                       void *stream);
int bls_reset_schur_sparse_mixed(bls_schur *h, double *nzval, const float *R_nzval, const double *scaling_r,
                                 double lambda, const double *s_r, void *stream);
int bls_reset_schur_dense_mixed(bls_schur *h, double *S, const float *R, const double *scaling_r, double lambda,
                                const double *s_r, void *stream);
int bls_reduce_schur_sparse_mixed(bls_schur *h, double *nzval, double *g_S, const double *b_r, const double *b_e,
                                  const double *scaling_e, double lambda, const float *E_data, const float *C_data,
                                  double *E_inv, double *z, const double *s_r, const double *s_e,
                                  int64_t *failed_block, void *stream);
int bls_reduce_schur_dense_mixed(bls_schur *h, double *S, double *g_S, const double *b_r, const double *b_e,
                                 const double *scaling_e, double lambda, const float *E_data, const float *C_data,
                                 double *E_inv, double *z, const double *s_r, const double *s_e,
                                 int64_t *failed_block, void *stream);
int bls_reduce_schur_sparse_mixed_async(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                        const double *b_e, const double *scaling_e, double lambda,
                                        const float *E_data, const float *C_data, double *E_inv, double *z,
                                        const double *s_r, const double *s_e, int64_t *d_failed_block,
                                        void *stream);
int bls_backsubstitute_schur_mixed(bls_schur *h, double *x_e, const double *x_r, const float *C_data,
                                   const double *E_inv, const double *z, const double *s_r, const double *s_e,
                                   void *stream);

/* A device copy of a scatter map (1-based Int64, host) for repeated copyto!: returns a device pointer. */
int bls_scatter_map_create(int64_t n, const int64_t *scatter_map, int64_t **d_map);

/* The shape of the phase-2 work list (no reference counterpart; the reference's col_chunks are sized by its own
 * work target, src/schur_reduction.jl:30-49): `part` landmarks per workgroup before a run of block rows of one column
 * is split into parts summed in part order (default 384), `lanes` scalar rows of one run before the next block row
 * starts another (default 96). Rebuilds the work list. Every shape sums in a fixed order of its own: deterministic;
 * shapes differ from each other by rounding only. */
int bls_schur_set_tuning(bls_schur *h, int64_t part, int64_t lanes);
int bls_schur_get_tuning(bls_schur *h, int64_t *part, int64_t *lanes);
/* Which phase-2 kernel a sparse handle runs (no reference counterpart; the reference has one loop nest,
 * src/schur_reduction.jl:315-380). mode 0 (default): the 16 x 16 tiles of S on the matrix cores when the layout fits
 * them (every landmark block of one size d_e <= 4, every retained block of one size 6..16, the default storage
 * offsets, each landmark coupling a contiguous run of retained blocks), else the work list; 1: always the work list;
 * 2: always the tiles (BLS_ERR_UNSUPPORTED when the layout does not fit). The kernels sum in different fixed orders:
 * each is deterministic, the two differ by rounding only. */
int bls_schur_set_phase2_mode(bls_schur *h, int mode);
/* info[0]: the phase-2 kernel in use (1 work list, 2 tiles); info[1..3]: the tile plan's tiles, landmark entries and
 * matrix-core products (0 when the layout does not fit). */
int bls_schur_phase2_info(bls_schur *h, int64_t *info);
/* The autotuner: times three Float64 reductions (the arguments of bls_reduce_schur_{sparse,dense}_f64, on the
 * handle's layout) for each of the ncand (parts[i], lanes[i]) shapes, after one untimed run each, and keeps the shape
 * with the fastest median (*best_ms: that median). In phase-2 mode 0 it also times the tiles when the layout fits
 * them and sets mode 1 when a work-list shape was faster. S is reduced repeatedly without a reset, so S, g_S, E_inv and z are scratch for the timing:
 * reset and reduce again after tuning. Synchronizes `stream`. */
int bls_schur_autotune_f64(bls_schur *h, int64_t ncand, const int64_t *parts, const int64_t *lanes, double *S,
                           double *g_S, const double *b_r, const double *b_e, const double *scaling_e, double lambda,
                           const double *E_data, const double *C_data, double *E_inv, double *z, const double *s_r,
                           const double *s_e, float *best_ms, void *stream);

/* Optional per-kernel timing of the last call on a handle (HIP events on the launch stream), in ms:
 * times[0] reset, [1] phase 1, [2] phase 2 (k_phase2, or k_tiles), [3] the merge of split block columns. Enabled by
 * bls_schur_set_timing(h, 1). */
int bls_schur_set_timing(bls_schur *h, int enable);
int bls_schur_kernel_times(bls_schur *h, float *times, int n);

/* ---- QR path (the Jacobian form) ---------------------------------------------------------------------- */

/* Register a SchurJacobianStructure (src/schur_jacobian.jl:30-57): the eliminated blocks' sizes and offsets
 * (e_dims, boffsets_e of its SchurStructure) and the table layout (table_nrows, table_ncols, raw_offsets,
 * residual_offsets; ntables + 1 offsets). Eliminated tables of at most 64 rows and d_e <= 8 take the register
 * kernel (four tables per wavefront); longer tables and d_e up to 4096 take one workgroup each (d_e above 4096:
 * BLS_ERR_UNSUPPORTED). */
int bls_qr_create(int64_t nr, int64_t nrblocks, int64_t neblocks, const int64_t *e_dims, const int64_t *boffsets_e,
                  int64_t ntables, const int64_t *table_nrows, const int64_t *table_ncols, const int64_t *raw_offsets,
                  const int64_t *residual_offsets, bls_qr **out);
void bls_qr_destroy(bls_qr *h);
/* Filling the tables from per-row-block Jacobian blocks: for every row block k of a structure built from a
 * description, block(J, k, blockentry(jst, k, j)) .= J_kj for each column block j it touches and
 * residual_block(J, k) .= r_k (src/schur_jacobian.jl:330-387; the views coupling_block / eliminated_block /
 * table_residual_block of :264-320). bls_jfill_create registers the row blocks (row_table, row_first, row_nrows,
 * row_ptr (nrowblocks + 1), row_blocks: the layout column blocks of each row block ascending, retained 1..nrblocks
 * and eliminated nrblocks + e; all 1-based, as SchurJacobianStructure(blocksizes, rowblocks, rowblocksizes;
 * eliminated) lays them out, src/schur_description.jl:137-189) against the table layout, and returns in *nvals and
 * *nres the lengths of the value and residual arrays a fill takes: row block after row block, each row block's
 * blocks in ascending column block, m_k x d column-major; its residual rows likewise. A row block that touches a
 * block its table does not hold, or rows outside its table's data rows, is BLS_ERR_ARG with the reference's reason.
 * bls_jfill_{f64,f32} write those values into the raw tables (element type T) and the Float64 residual, asynchronously
 * on the stream; the other rows (damping rows, rows of other row blocks) are left as they are. */
int bls_jfill_create(int64_t nrblocks, int64_t neblocks, const int64_t *r_dims, const int64_t *e_dims, int64_t ntables,
                     const int64_t *table_nrows, const int64_t *table_ncols, const int64_t *table_top,
                     const int64_t *raw_offsets, const int64_t *residual_offsets, const int64_t *table_ptr,
                     const int64_t *entry_retained, const int64_t *entry_col, int64_t nrowblocks,
                     const int64_t *row_table, const int64_t *row_first, const int64_t *row_nrows,
                     const int64_t *row_ptr, const int64_t *row_blocks, int64_t *nvals, int64_t *nres,
                     bls_jfill **out);
void bls_jfill_destroy(bls_jfill *h);
int bls_jfill_f64(bls_jfill *h, double *raw, double *r, const double *vals, const double *res, void *stream);
int bls_jfill_f32(bls_jfill *h, float *raw, double *r, const float *vals, const double *res, void *stream);
/* reduce_qr!(ws, J, scaling_e, lambda, s) (src/qr_reduction.jl:162-187): factored = Q' [J_r | J_e] per table with
 * the damping rows sqrt(lambda scaling_e) written, fres = Q' r in Float64, free tables copied; raw tables and r
 * are read only. BLS_FAILED with the lowest failing eliminated block for a non-finite table or a numerically
 * singular U (|beta| <= M eps ||col||). `s` is the conditioning scale over all nr + ne variables (device
 * Float64, NULL = UnitScale; scale_table_columns!, src/qr_reduction.jl:236-255): each column is scaled by its
 * variable's s as it is read and the damping rows by the eliminated variable's; it needs bls_qr_set_reduced. */
int bls_reduce_qr_f64(bls_qr *h, double *factored, double *fres, const double *raw, const double *r,
                      const double *scaling_e, double lambda, const double *s, int64_t *failed_block, void *stream);
int bls_reduce_qr_f64_async(bls_qr *h, double *factored, double *fres, const double *raw, const double *r,
                            const double *scaling_e, double lambda, const double *s, int64_t *d_failed_block,
                            void *stream);
/* reduce_qr! of a SchurBlockJacobian{Float32} (src/qr_linear_systems.jl:2-14: tables and their elimination in T,
 * the residual in Float64): raw and factored tables in Float32, fres in Float64. The reflectors are formed and
 * applied in Float64 registers from the Float32 tables and rounded to Float32 once; the singularity test uses
 * eps(Float32) as the reference's does. */
int bls_reduce_qr_f32(bls_qr *h, float *factored, double *fres, const float *raw, const double *r,
                      const double *scaling_e, double lambda, const double *s, int64_t *failed_block, void *stream);
int bls_reduce_qr_f32_async(bls_qr *h, float *factored, double *fres, const float *raw, const double *r,
                            const double *scaling_e, double lambda, const double *s, int64_t *d_failed_block,
                            void *stream);
/* The table entries and reduced-Jacobian structure of the registered SchurJacobianStructure
 * (src/schur_jacobian.jl:30-160), for the products below and for reduce_qr's conditioning scale: table_top
 * (ntables), reduced_offsets (ntables + 1), table_ptr (ntables + 1), and per table
 * entry entry_retained, entry_col, entry_table (nentries), retained_ptr (nrblocks + 1), retained_entry (nentries),
 * r_dims / boffsets_r of its SchurStructure, doffsets (nrblocks + 1). All 1-based as stored. */
int bls_qr_set_reduced(bls_qr *h, const int64_t *table_top, const int64_t *reduced_offsets, const int64_t *table_ptr,
                       int64_t nentries, const int64_t *entry_retained, const int64_t *entry_col,
                       const int64_t *entry_table, const int64_t *retained_ptr, const int64_t *retained_entry,
                       const int64_t *r_dims, const int64_t *boffsets_r, const int64_t *doffsets);
/* The sparse reduced-system layout for reduced_normal_sparse!: S.colptr (nr + 1) and the block pattern of
 * build_schur_pattern (src/schur_structure.jl:241-323: block_colptr, block_rowval, block_row_offsets, col_heights),
 * with the structure's table_ptr / entry_retained. Every pair of entries of a table must have its block in the
 * pattern (BLS_ERR_ARG otherwise: the reference's throw_missing_block). Needs bls_qr_set_reduced. */
int bls_qr_set_sparse(bls_qr *h, int64_t nnz, const int64_t *S_colptr, int64_t nnz_blocks, const int64_t *block_colptr,
                      const int64_t *block_rowval, const int64_t *block_row_offsets, const int64_t *col_heights,
                      const int64_t *table_ptr, const int64_t *entry_retained);
/* reduced_normal_sparse!(S, g_S, ws, pattern) / reduced_normal_dense!(S, g_S, ws) (src/qr_reduced.jl:373-480):
 * every scalar column of S is zeroed, then its upper triangle is J~'J~ and g_S = J~'r~ (Float64), each S entry summed
 * over the tables in the reference's order (deterministic). Dense S is nr x nr column-major. A table may hold up to
 * 64 retained blocks spanning up to 1024 columns (BLS_ERR_UNSUPPORTED otherwise), in strictly ascending order
 * (BLS_ERR_ARG otherwise, as the reference's structure builds them). The first call per layout builds a plan and
 * allocates a Float64 contribution buffer of sum over tables of sum over entry pairs of d1 d2 values. */
int bls_reduced_normal_sparse_f64(bls_qr *h, double *S_nzval, double *g_S, const double *factored, const double *fres,
                                  void *stream);
int bls_reduced_normal_sparse_f32(bls_qr *h, float *S_nzval, double *g_S, const float *factored, const double *fres,
                                  void *stream);
int bls_reduced_normal_dense_f64(bls_qr *h, double *S, double *g_S, const double *factored, const double *fres,
                                 void *stream);
int bls_reduced_normal_dense_f32(bls_qr *h, float *S, double *g_S, const float *factored, const double *fres,
                                 void *stream);
/* backsubstitute!(x_e, x_r, ws::QRReductionWorkspace) (src/qr_reduced.jl:24-110): x_e = U^-1 (q - B x_r) per
 * eliminated table, from the factored tables and residual of the last reduce_qr; accumulated in Float64. Needs
 * bls_qr_set_reduced. */
int bls_qr_backsubstitute_f64(bls_qr *h, double *x_e, const double *x_r, const double *factored, const double *fres,
                              void *stream);
int bls_qr_backsubstitute_f32(bls_qr *h, double *x_e, const double *x_r, const float *factored, const double *fres,
                              void *stream);
/* reduced_mul! (src/qr_reduced.jl:161-193): y = J~ v, J~ the rows of the factored tables below U. */
int bls_reduced_mul_f64(bls_qr *h, double *y, const double *factored, const double *v, void *stream);
int bls_reduced_mul_f32(bls_qr *h, double *y, const float *factored, const double *v, void *stream);
/* reduced_tmul! (src/qr_reduced.jl:195-223): g = J~' w (deterministic: fixed-order partial sums). */
int bls_reduced_tmul_f64(bls_qr *h, double *g, const double *factored, const double *w, void *stream);
int bls_reduced_tmul_f32(bls_qr *h, double *g, const float *factored, const double *w, void *stream);
/* reduced_residual! (src/qr_reduced.jl:227-243): r~, the factored residual rows below U. */
int bls_reduced_residual_f64(bls_qr *h, double *rt, const double *fres, void *stream);
/* Quantities of the full problem from the tables F (the raw tables J.data, or the factored ws.factored) over all
 * rows, accumulated in Float64 (src/qr_reduced.jl:574-645); needs bls_qr_set_reduced. Vectors are [v_r; v_e] of
 * length nr + sum of e_dims.
 * jtr! (:574-583): g = J'r, r the Float64 residual at residual_offsets (J.r, or ws.fres).
 * jtj_diag! (:585-595): d = diag(J'J), the squared column norms (colnorms2!, src/qr_linear_systems.jl:58-64).
 * jtj_quadform (:597-645): ||J v||^2 into the device double *result, the tables' partials added in a fixed order.
 * A retained column's value adds its tables' in retained_entry order, in runs (deterministic). */
int bls_jtr_f64(bls_qr *h, double *g, const double *F, const double *r, void *stream);
int bls_jtr_f32(bls_qr *h, double *g, const float *F, const double *r, void *stream);
int bls_jtj_diag_f64(bls_qr *h, double *d, const double *F, void *stream);
int bls_jtj_diag_f32(bls_qr *h, double *d, const float *F, void *stream);
int bls_jtj_quadform_f64(bls_qr *h, const double *F, const double *v, double *result, void *stream);
int bls_jtj_quadform_f32(bls_qr *h, const float *F, const double *v, double *result, void *stream);
/* build_reduced_jacobian_pattern (src/qr_reduced.jl:289-310) and reduced_jacobian! (:247-270, :334-370): J~ (the
 * reduced rows of every table's retained columns, reduced_nrows x nr) written out for a sparse or dense solver.
 * bls_reduced_jacobian_nnz: the pattern's stored entries. bls_build_reduced_jacobian_pattern: its host colptr (nr + 1)
 * and rowval (nnz), 1-based: each column of retained block r holds the reduced rows of r's tables, tables in
 * retained_entry order. bls_check_reduced_jacobian_pattern (check_reduced_jacobian_pattern, :312-332): a host CSC of
 * size m x n against that pattern (size, column heights, the first row of every run), with the reference's errors;
 * the caller checks a matrix once before filling it. The fills write J~ from the factored tables F (sparse: into the
 * nzval of that pattern; dense: a column-major m x nr array, zeroed first) and rt = r~ (Float64) from fres. */
int bls_reduced_jacobian_nnz(bls_qr *h, int64_t *nnz);
int bls_build_reduced_jacobian_pattern(bls_qr *h, int64_t *colptr, int64_t *rowval);
int bls_check_reduced_jacobian_pattern(bls_qr *h, int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval);
int bls_reduced_jacobian_sparse_f64(bls_qr *h, double *nzval, double *rt, const double *F, const double *fres,
                                    void *stream);
int bls_reduced_jacobian_sparse_f32(bls_qr *h, float *nzval, double *rt, const float *F, const double *fres,
                                    void *stream);
int bls_reduced_jacobian_dense_f64(bls_qr *h, double *Jt, double *rt, const double *F, const double *fres, void *stream);
int bls_reduced_jacobian_dense_f32(bls_qr *h, float *Jt, double *rt, const float *F, const double *fres, void *stream);
/* reduced_diagonal_blocks! (src/qr_reduced.jl:484-512): the diagonal blocks of J~'J~, full d x d column-major at
 * doffsets (the block-Jacobi blocks of the QR path). */
int bls_reduced_diagonal_blocks_f64(bls_qr *h, double *D, const double *factored, void *stream);
int bls_reduced_diagonal_blocks_f32(bls_qr *h, float *D, const float *factored, void *stream);
/* Kernel time of the last synchronous reduce_qr (HIP events, ms), when enabled. */
int bls_qr_set_timing(bls_qr *h, int enable);
int bls_qr_kernel_time(bls_qr *h, float *ms);

/* ---- block-Jacobi PCG on the reduced system (no reference counterpart) -------------------------------- */

typedef struct bls_pcg bls_pcg;
/* A PCG solver for reduced systems in the layout of build_schur_pattern (S_colptr, block pattern; as for
 * bls_schur_set_sparse), batched over `nprob` independent problems whose retained blocks are the 1-based ranges
 * prob_block_ptr[p] .. prob_block_ptr[p+1]-1 (no block of S may couple two problems). The preconditioner is the
 * reference's BlockJacobi (ext/KrylovExt.jl:88-95: factor_diagonal_blocks! / solve_diagonal_blocks!,
 * src/diagonal_blocks.jl:6-61) on S's diagonal blocks. */
int bls_pcg_create(int64_t nr, int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r, int64_t nnz,
                   const int64_t *S_colptr, int64_t nnz_blocks, const int64_t *block_colptr,
                   const int64_t *block_rowval, const int64_t *block_row_offsets, const int64_t *col_heights,
                   int64_t nprob, const int64_t *prob_block_ptr, bls_pcg **out);
void bls_pcg_destroy(bls_pcg *h);
/* Sizes of the handle's symmetric block-row copy of S and of the preconditioner blocks (sum d_r^2). */
int bls_pcg_values_size(bls_pcg *h, int64_t *nvals, int64_t *nfactor);
/* Solve S x = g (device arrays; S in the upper-block CSC of build_schur_pattern) per problem until the true
 * residual ||S x - g|| / ||g|| <= tol (FP64), or maxit iterations. precond_blocks: NULL for the plain
 * block-Jacobi preconditioner, else full d x d column-major SPD blocks (sum d_r^2 values) that replace S's
 * diagonal blocks in it (the learned correction). Writes iters[p] (-1: preconditioner block not positive
 * definite) and relres[p] (host arrays of nprob). BLS_OK when every problem converged, else BLS_FAILED. */
int bls_pcg_solve_f64(bls_pcg *h, const double *S_nzval, const double *g, double *x, double tol, int64_t maxit,
                      const double *precond_blocks, int64_t *iters, double *relres, void *stream);
/* y += alpha x (device vectors of n doubles): the update of an iterative refinement step. */
int bls_axpy_f64(int64_t n, double alpha, const double *x, double *y, void *stream);
/* r = t - S x, and per problem ||r|| / ||gref|| into the host array relres (nprob values); synchronizes. With S the
 * reset S (R + damping, bls_reset_schur_sparse_f64), t = bls_schur_retained_residual_f64(x_e = backsubstitute!(x))
 * and gref = g_S, r is the reduced system's residual g_S - S_red x evaluated from the blocks, without S_red. No
 * reference counterpart (the Float64 iterative refinement of the mixed reduction). */
int bls_pcg_residual_f64(bls_pcg *h, const double *S_nzval, const double *t, const double *x, const double *gref,
                         double *r, double *relres, void *stream);
/* The true residual ||S x - g|| / ||g|| of each problem for a given x (device arrays; relres: host, nprob), e.g.
 * of one system's solution against another system of the same pattern. Synchronizes `stream`. */
int bls_pcg_relres_f64(bls_pcg *h, const double *S_nzval, const double *g, const double *x, double *relres,
                       void *stream);

/* ---- block-diagonal kernels (src/diagonal_blocks.jl:6-61) ---------------------------------------------- */

typedef struct bls_diag bls_diag;   /* the block sizes of a flat array of square blocks */

/* Register `dims` (nblocks sizes, 1..16): block b is column-major at the sum of the d^2 before it. */
int bls_diag_create(int64_t nblocks, const int64_t *dims, bls_diag **out);
void bls_diag_destroy(bls_diag *h);
/* factor_diagonal_blocks!(D, dims) (src/diagonal_blocks.jl:6-30): the Cholesky factor of every block in its lower
 * triangle, in the element type. BLS_FAILED when a block is not positive definite, with the lowest such block
 * (1-based) in *failed_block; the reference stops at that block, the device factors the others too, and D is then
 * not to be used, as the reference's caller treats it. Synchronizes `stream`. */
int bls_factor_diagonal_blocks_f64(bls_diag *h, double *D, int64_t *failed_block, void *stream);
int bls_factor_diagonal_blocks_f32(bls_diag *h, float *D, int64_t *failed_block, void *stream);
/* solve_diagonal_blocks!(y, D, dims, x) (src/diagonal_blocks.jl:33-61): y = (L L')^-1 x per block, in Float64;
 * y may be x. */
int bls_solve_diagonal_blocks_f64(bls_diag *h, double *y, const double *D, const double *x, void *stream);
int bls_solve_diagonal_blocks_f32(bls_diag *h, double *y, const float *D, const double *x, void *stream);

/* ---- learned correction of the block-Jacobi preconditioner (no reference counterpart) ------------------ */

/* Per retained block of a PCG handle's reduced system (device S in build_schur_pattern's layout), 26 FP32
 * features (nrblocks x 26, row-major) for the MLP: the 21 upper entries of the diagonal block D divided by the mean
 * m of its diagonal, log10 m, log(1 + off-diagonal blocks in the block row), log10 lambda, log10(m / lambda),
 * log10(off-diagonal Frobenius mass of the block row / ||D||_F). Blocks of dimension other than 6 get zeros. */
int bls_pcg_features_f32(bls_pcg *h, const double *S_nzval, double lambda, float *features, void *stream);
/* The group correction of block-Jacobi (the learned correction's family; no reference counterpart).
 * bls_pcg_set_groups(h, G) groups the retained blocks of each problem in order, G at a time (G = 1..4; the last group
 * of a problem may be shorter). The preconditioner block of a group is S's principal block submatrix over its
 * blocks with every off-diagonal block multiplied by w = omega[b] in [0, 1] (b the group's first block): w = 0 is
 * plain block-Jacobi, w = 1 the group's exact block of S, SPD for every such w. bls_pcg_group_blocks_size gives the
 * length of the group blocks, bls_pcg_group_blocks_f64 forms them from S (omega: Float32 per retained block), and
 * bls_pcg_solve_groups_f64 is bls_pcg_solve_f64 with them as the preconditioner. */
int bls_pcg_set_groups(bls_pcg *h, int group);
int bls_pcg_group_blocks_size(bls_pcg *h, int64_t *n);
int bls_pcg_group_blocks_f64(bls_pcg *h, const double *S_nzval, const float *omega, double *blocks, void *stream);
int bls_pcg_solve_groups_f64(bls_pcg *h, const double *S_nzval, const double *g, double *x, double tol, int64_t maxit,
                             const double *group_blocks, int64_t *iters, double *relres, void *stream);
/* The corrected preconditioner blocks M_b = D_b + mu_b m_b I (mu: nrblocks FP32 shifts, >= 0 keeps M_b SPD when D_b
 * is), full d x d column-major (sum d^2 values): the precond_blocks argument of bls_pcg_solve_f64. */
int bls_pcg_shift_blocks_f64(bls_pcg *h, const double *S_nzval, const float *mu, double *blocks, void *stream);

/* The MLP (26 -> 64 -> 64 -> 1, ReLU, FP32): sizes of its feature vector, hidden layers, flat parameter vector
 * (W1 [26][64], b1, W2 [64][64], b2, W3 [64], b3) and per-sample work space of bls_mlp_grad_f32. */
int bls_mlp_dims(int *nfeatures, int *nhidden, int *nparams, int *work_per_sample);
/* out[s] = MLP((X[s] - fmean) / fstd) for n samples (X: n x 26 row-major). */
int bls_mlp_forward_f32(const float *params, const float *fmean, const float *fstd, const float *X, int64_t n,
                        float *out, void *stream);
/* Outputs and the gradient of L = mean_s (out[s] - y[s])^2 / 2 with respect to the parameters (deterministic:
 * every gradient entry is summed in sample order). work: n * work_per_sample floats of scratch. */
int bls_mlp_grad_f32(const float *params, const float *fmean, const float *fstd, const float *X, const float *y,
                     int64_t n, float *out, float *grad, float *work, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected for step >= 1) on the parameters. */
int bls_mlp_adam_f32(float *params, const float *grad, float *m, float *v, int64_t step, float lr, void *stream);

/* ---- assembly of a Schur Hessian from row blocks (src/schur_assembly.jl:140-246) --------------------------- */

typedef struct bls_asm bls_asm;   /* the plan of one fill pattern */

/* Register the fill pattern that addblock!(H, i, j, J_k'J_k, rows_i, cols_j) (src/schur_assembly.jl:140-159) would
 * follow for a batch of row blocks, every pair of blocks of each row block, row block by row block. Structure: the
 * SchurStructure fields as for bls_schur_create. R: r_colptr == NULL for a dense nr x nr column-major Matrix (upper
 * triangle and full diagonal blocks written, as add_retained_diagonal! / add_retained_pair! do, :189-223), else a
 * SparseBlockLinearSystem's stored lower blocks (src/SparseBlockLinearSystem.jl:11-64): r_colptr (nrblocks + 1),
 * r_rowval (the block rows >= each block column, ascending), r_offsets (start of each block in R.data, column-major
 * d_row x d_col), all 1-based. Row blocks: rb_ptr (nrowblocks + 1) into rb_blocks, the layout block numbers
 * (1-based: retained 1..nrblocks, eliminated nrblocks + e, as addblock! numbers them) in the column order of the
 * row block's Jacobian, and rb_rows its row count. A pair the system has no storage for is BLS_ERR_ARG with the
 * reference's reason (check_addblock, :161-181); a row block over more than 64 blocks or 1024 columns is
 * BLS_ERR_UNSUPPORTED. */
int bls_asm_create(int64_t nrblocks, int64_t neblocks, const int64_t *r_dims, const int64_t *e_dims,
                   const int64_t *boffsets_r, const int64_t *boffsets_e, const int64_t *eoffsets,
                   const int64_t *eliminated_ptr, const int64_t *entry_retained, const int64_t *coffsets,
                   const int64_t *r_colptr, const int64_t *r_rowval, const int64_t *r_offsets, int64_t nrowblocks,
                   const int64_t *rb_ptr, const int64_t *rb_blocks, const int64_t *rb_rows, bls_asm **out);
void bls_asm_destroy(bls_asm *a);
/* Lengths the arrays of bls_asm_hessian_f64 need: E_data, C_data, R, and nr. */
int bls_asm_lengths(const bls_asm *a, int64_t *lengths /* 4 */);
/* H = sum_k J_k' J_k into E_data, C_data and R, and b = sum_k J_k' r_k into b_r and b_e: what a SchurBlockHessian
 * that starts at zero holds after the addblock! calls of the registered pattern, made in row-block order (each
 * stored entry summed over the row blocks in that order; bitwise reproducible). Every stored block is written
 * (zero where no row block touches it; a dense R is zeroed first). J: the row blocks one after another, each
 * rb_rows x (sum of its block dims) column-major; r: their residual rows one after another. */
int bls_asm_hessian_f64(bls_asm *a, const double *J, const double *r, double *E_data, double *C_data, double *R,
                        double *b_r, double *b_e, void *stream);
/* dot(x, H, y) (src/linalg.jl:133-260; schur_quadform when x == y): x' H y over the storage of the registered
 * layout (a handle registered with nrowblocks = 0 serves for this alone), x and y laid out [x_r; x_e], length
 * nr + sum of e_dims. A dense R is read in its upper triangle, a sparse R by its stored blocks, each off-diagonal
 * block both ways. The terms are added in a fixed order (bitwise reproducible); the sum is written to the device
 * double *result when the stream reaches it. */
int bls_asm_dot_f64(bls_asm *a, const double *x, const double *y, const double *E_data, const double *C_data,
                    const double *R, double *result, void *stream);
/* The same over a SchurBlockHessian{Float32}'s storage, each stored value widened and the terms added in Float64
 * (the promoted type of Float32 blocks and Float64 vectors, src/linalg.jl:12). */
int bls_asm_dot_f32(bls_asm *a, const double *x, const double *y, const float *E_data, const float *C_data,
                    const float *R, double *result, void *stream);
/* getdiag! (src/schur_linear_systems.jl:239-270): d = [diag(R); diag(E_1); ...], the undamped diagonal of H as
 * filled, length nr + sum of e_dims; a retained block whose diagonal block a sparse R does not store reads zeros. */
int bls_asm_diag_f64(bls_asm *a, const double *E_data, const double *R, double *d, void *stream);
/* The same from a SchurBlockHessian{Float32}'s storage into the Float64 d (EliminationConditioning.d). */
int bls_asm_diag_f32(bls_asm *a, const float *E_data, const float *R, double *d, void *stream);

/* ---- The dense reduced solve (src/schur_linear_systems.jl:312-319) ----------------------------------------
 * The :cholesky branch of solve_reduced_dense!: S = U'U for the dense reduced system S (n x n column-major, its
 * upper triangle as reduce_schur_dense! writes it; the lower triangle is not read), factored in a copy held by the
 * handle (S is left as built), then x = S^-1 g. The factorization is DampedLinearSolve's (LAPACK potrf): a pivot that
 * is not positive (or NaN) stops it, BLS_FAILED with the 1-based column in *failed_col (NotPositiveDefinite: the
 * caller falls back to an indefinite factorization). */
typedef struct bls_chol bls_chol;
int bls_chol_create(int64_t n, bls_chol **out);
void bls_chol_destroy(bls_chol *h);
int bls_chol_factor_f64(bls_chol *h, const double *S, int64_t *failed_col, void *stream);
/* The same, queued without waiting: a failure is then seen by bls_chol_solve_f64 (which leaves x as g) and by the
 * next bls_chol_factor_f64. */
int bls_chol_factor_f64_async(bls_chol *h, const double *S, void *stream);
/* A Float32 S (a SchurBlockHessian{Float32}'s reduced system) is widened into the handle's Float64 copy and factored
 * in Float64. */
int bls_chol_factor_f32(bls_chol *h, const float *S, int64_t *failed_col, void *stream);
int bls_chol_factor_f32_async(bls_chol *h, const float *S, void *stream);
/* r = t - S x for a symmetric S (n x n, upper triangle stored column-major, as reset_schur_dense! writes it): the
 * product with R + damping in the dense refinement residual. No reference counterpart. */
int bls_dense_residual_f64(int64_t n, const double *S, const double *t, const double *x, double *r, void *stream);
int bls_chol_solve_f64(bls_chol *h, const double *g, double *x, void *stream);
/* The factor U (upper triangle of an n x n column-major array; the lower triangle is unspecified), for tests. */
int bls_chol_factor_copy_f64(bls_chol *h, double *U, void *stream);

/* The :ldl branch of solve_reduced_dense! (src/schur_linear_systems.jl:312-319): Bunch-Kaufman of a copy of the
 * dense symmetric S (its upper triangle read, n x n column-major), the fallback when Cholesky fails. The reference
 * reaches it through DampedLinearSolve's bunchkaufman (LAPACK sytrf); this is LAPACK's unblocked dsytf2 on the lower
 * triangle of the copy (1 x 1 and 2 x 2 pivots, symmetric interchanges, LAPACK's ipiv). bls_ldl_factor_f64
 * synchronizes: BLS_OK with *negative the number of negative eigenvalues of D (those of S; > 0 is the reference's
 * Indefinite status), or BLS_FAILED with *zero_col the first zero pivot (1-based; S singular). bls_ldl_solve_f64:
 * x = S^-1 g with the last factor (dsytrs), queued on the stream; x may be g. bls_ldl_factor_copy_f64: the factor
 * (L below the diagonal, D on it and, for a 2 x 2 pivot, below it) and ipiv, device to device. */
typedef struct bls_ldl bls_ldl;
int bls_ldl_create(int64_t n, bls_ldl **out);
void bls_ldl_destroy(bls_ldl *h);
int bls_ldl_factor_f64(bls_ldl *h, const double *S, int64_t *zero_col, int64_t *negative, void *stream);
int bls_ldl_solve_f64(bls_ldl *h, const double *g, double *x, void *stream);
int bls_ldl_factor_copy_f64(bls_ldl *h, double *A, int64_t *ipiv, void *stream);

/* ---- The conditioning scale (src/conditioning.jl:24-87) ----------------------------------------------------
 * bls_cond holds EliminationConditioning's state (:47-55) except its vectors: whether the next prepare computes the
 * scale afresh (stale; set at creation and by bls_cond_markstale, as zero! and markstale! set it) and the number of
 * recomputations at a later damping (nrescale). The caller owns s and d (n device doubles each). */
typedef struct bls_cond bls_cond;
int bls_cond_create(int64_t n, bls_cond **out);
void bls_cond_destroy(bls_cond *c);
/* 1 if the next bls_prepare_scale_f64 computes the scale from d: the caller's diag! (bls_asm_diag_f64 on the
 * Hessian side, bls_jtj_diag_* on the Jacobian side) writes d first; 0 if it reuses the scale; < 0 on error. */
int bls_cond_is_stale(const bls_cond *c);
int bls_cond_markstale(bls_cond *c);
/* prepare_scale!(c, diag!, lambda, sc, T) (:60-87) with d already written when stale: s_j = conditioning_scale(d_j +
 * max(lambda, 0) sc_j) (:29-35; sc NULL = ones) for every j when stale; otherwise the same for every j if, for some
 * variable, |d_j + lambda sc_j| s_j^2 lies outside [2^-16, 2^16] and its scale would change (nrescale += 1), else s is
 * left as it is -- decided on the device. elem_f32: the system's element type is Float32 (the scale's floor is then
 * floatmin(Float32)). Clears stale. */
int bls_prepare_scale_f64(bls_cond *c, double *s, const double *d, const double *sc, double lambda, int elem_f32,
                          void *stream);
/* nrescale, after synchronizing the stream. */
int bls_cond_nrescale(bls_cond *c, int64_t *out, void *stream);
/* unscale!(x, s) (:90-95) for the n variables of the handle: x = S y in place (s NULL = UnitScale: x untouched), and
 * the device int *nonfinite = 1 if x then holds a value that is not finite (the reference returns
 * NotPositiveDefinite then, src/schur_linear_systems.jl:333-334), else 0. */
int bls_unscale_f64(bls_cond *c, double *x, const double *s, int32_t *nonfinite, void *stream);

#ifdef __cplusplus
}
#endif
#endif
