/*
 * bls_oracle_t.h -- element-type template of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 * Included twice by bls_oracle.c, with T/S = double/f64 and float/f32. Indices are the reference's 1-based
 * values; the macros below turn them into C offsets at the point of use, so every formula reads as in Julia.
 */
#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, S)

/* 1-based accessors: A1(v, i) is v[i] in Julia */
#define A1(v, i) ((v)[(i) - 1])

/* scaled(T, a, f) = T(Float64(a) * f) (src/conditioning.jl:202); NULL scale = UnitScale (exactly a) */
static inline T FN(scaled)(T a, double f) { return (T)((double)a * f); }
#define SC(v, i) ((v) ? A1(v, i) : 1.0)

/* ---------------------------------------------------------------------------------------------------------
 * inv_sym_block (src/kernels.jl:14-83): damped SPD inverse with the leading-minor / pivot-ratio test.
 * H is d x d column-major (already scaled); damp[i] is the damping diagonal (lambda * scale, already in T);
 * the closed forms add it as H_ii + 1*damp_i exactly as `H[i,i] + lambda*scale[i]` with lambda = one(T).
 * Returns 1 on success.
 * ------------------------------------------------------------------------------------------------------- */
int FN(orc_inv_sym_block)(int64_t d, const T *H, const T *damp, T *invH)
{
    const T tol = (T)4504 * (sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07); /* kernels.jl:3 */
    if (d == 1) { /* kernels.jl:14-21 */
        T h11 = H[0] + damp[0];
        if (!(h11 > 0) || !isfinite(h11)) { invH[0] = 0; return 0; }
        invH[0] = (T)1 / h11;
        return 1;
    }
    if (d == 2) { /* kernels.jl:23-36 */
        T H11 = H[0] + damp[0], H22 = H[3] + damp[1], H12 = H[2];
        T det = H11 * H22 - H12 * H12;
        if (!(H11 > 0) || !(det > tol * (H11 * H22)) || !isfinite(det)) {
            for (int i = 0; i < 4; i++) invH[i] = 0;
            return 0;
        }
        T inv_det = (T)1 / det;
        invH[0] = H22 * inv_det; invH[1] = -H12 * inv_det; invH[2] = -H12 * inv_det; invH[3] = H11 * inv_det;
        return 1;
    }
    if (d == 3) { /* kernels.jl:38-67; H[i,j] column-major = H[(j-1)*3 + i-1] */
        T H11 = H[0] + damp[0], H22 = H[4] + damp[1], H33 = H[8] + damp[2];
        T H12 = H[3], H13 = H[6], H23 = H[7];
        T c11 = H22 * H33 - H23 * H23;
        T c12 = H13 * H23 - H12 * H33;
        T c13 = H12 * H23 - H13 * H22;
        T c22 = H11 * H33 - H13 * H13;
        T c23 = H12 * H13 - H11 * H23;
        T c33 = H11 * H22 - H12 * H12;
        T det = H11 * c11 + H12 * c12 + H13 * c13;
        if (!(H11 > 0) || !(c33 > tol * (H11 * H22)) || !(det > tol * (c33 * H33)) || !isfinite(det)) {
            for (int i = 0; i < 9; i++) invH[i] = 0;
            return 0;
        }
        T inv_det = (T)1 / det;
        invH[0] = c11 * inv_det; invH[1] = c12 * inv_det; invH[2] = c13 * inv_det;
        invH[3] = c12 * inv_det; invH[4] = c22 * inv_det; invH[5] = c23 * inv_det;
        invH[6] = c13 * inv_det; invH[7] = c23 * inv_det; invH[8] = c33 * inv_det;
        return 1;
    }
    return FN(orc_inv_sym_block_chol)(d, H, damp, invH);
}

/* The generic method of inv_sym_block (src/kernels.jl:69-83), callable at any size 1..8 (test/kernels.jl:11-49
 * calls it at sizes 1-3 to pin the closed forms): Cholesky of Symmetric(Hd, :U), every pivot ratio
 * |U_ii|^2 / Hd_ii > tol, inv = U^-1 U^-T. */
int FN(orc_inv_sym_block_chol)(int64_t d, const T *H, const T *damp, T *invH)
{
    const T tol = (T)4504 * (sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07);
    T U[64], Ui[64];
    if (d > 8) return 0;
    for (int64_t j = 0; j < d; j++)
        for (int64_t i = 0; i < d; i++) U[j * d + i] = 0;
    for (int64_t j = 0; j < d; j++) {
        for (int64_t i = 0; i <= j; i++) {
            T a = H[j * d + i] + (i == j ? damp[i] : (T)0);
            for (int64_t k = 0; k < i; k++) a -= U[i * d + k] * U[j * d + k];
            if (i == j) {
                if (!(a > 0) || !isfinite(a)) { for (int64_t q = 0; q < d * d; q++) invH[q] = 0; return 0; }
                U[j * d + j] = (T)sqrt((double)a);
                if (sizeof(T) == 4) U[j * d + j] = sqrtf((float)a);
            } else {
                U[j * d + i] = a / U[i * d + i];
            }
        }
    }
    T min_ratio = (T)INFINITY;
    for (int64_t i = 0; i < d; i++) {
        T hd = H[i * d + i] + damp[i];
        T r = (U[i * d + i] * U[i * d + i]) / hd;
        if (r < min_ratio || isnan(r)) min_ratio = r;
    }
    if (!(min_ratio > tol && isfinite(min_ratio))) { for (int64_t q = 0; q < d * d; q++) invH[q] = 0; return 0; }
    /* Ui = U^-1 (upper triangular) */
    for (int64_t q = 0; q < d * d; q++) Ui[q] = 0;
    for (int64_t j = 0; j < d; j++) {
        Ui[j * d + j] = (T)1 / U[j * d + j];
        for (int64_t i = j - 1; i >= 0; i--) {
            T s = 0;
            for (int64_t k = i + 1; k <= j; k++) s += U[k * d + i] * Ui[j * d + k];
            Ui[j * d + i] = -s / U[i * d + i];
        }
    }
    /* inv = Ui * Ui^T */
    for (int64_t j = 0; j < d; j++)
        for (int64_t i = 0; i < d; i++) {
            T s = 0;
            for (int64_t k = (i > j ? i : j); k < d; k++) s += Ui[k * d + i] * Ui[k * d + j];
            invH[j * d + i] = s;
        }
    return 1;
}

/* Cholesky (upper) + inverse of a d x d matrix with d > 8: the LAPACK potrf!/potri! route of
 * src/schur_reduction.jl:164-225, with the pivot ratio against the damped diagonal recomputed in double. */
static int FN(chol_inverse_dyn)(int64_t d, T *M, const double *hii)
{
    const double tol = 4504.0 * (sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-07);
    for (int64_t j = 0; j < d; j++) {
        for (int64_t i = 0; i <= j; i++) {
            T a = M[j * d + i];
            for (int64_t k = 0; k < i; k++) a -= M[i * d + k] * M[j * d + k];
            if (i == j) {
                if (!(a > 0)) return 0; /* potrf info != 0 */
                M[j * d + j] = (T)sqrt((double)a);
            } else {
                M[j * d + i] = a / M[i * d + i];
            }
        }
    }
    for (int64_t i = 0; i < d; i++) {
        double ratio = ((double)M[i * d + i] * (double)M[i * d + i]) / hii[i];
        if (!(isfinite(ratio) && ratio > tol)) return 0;
    }
    /* potri: inverse of U^T U from U, upper triangle */
    T *Ui = (T *)malloc(sizeof(T) * d * d);
    for (int64_t q = 0; q < d * d; q++) Ui[q] = 0;
    for (int64_t j = 0; j < d; j++) {
        Ui[j * d + j] = (T)1 / M[j * d + j];
        for (int64_t i = j - 1; i >= 0; i--) {
            T s = 0;
            for (int64_t k = i + 1; k <= j; k++) s += M[k * d + i] * Ui[j * d + k];
            Ui[j * d + i] = -s / M[i * d + i];
        }
    }
    for (int64_t j = 0; j < d; j++)
        for (int64_t i = 0; i <= j; i++) {
            T s = 0;
            for (int64_t k = j; k < d; k++) s += Ui[k * d + i] * Ui[k * d + j];
            M[j * d + i] = s;
            M[i * d + j] = s;
        }
    free(Ui);
    return 1;
}

/* ---------------------------------------------------------------------------------------------------------
 * Phase 1: factor_eliminated! (src/schur_reduction.jl:121-225) for eliminated block e (1-based).
 * ------------------------------------------------------------------------------------------------------- */
static int FN(factor_eliminated)(const orc_schur *st, T *E_inv, double *z, T *W, const T *E_data, const T *C_data,
                                 const double *b_e, const double *scaling_e, double lambda, int64_t e,
                                 const double *s_r, const double *s_e)
{
    const int64_t de = A1(st->e_dims, e);
    const int64_t eo = A1(st->eoffsets, e);
    const int64_t sc_off = A1(st->boffsets_e, e);
    if (de <= 8) {
        /* static path: damp = T.((lambda .* scaling) .* (se .* se)), Es[i,j] = scaled(E, se_i se_j) */
        T Es[64], damp[8], inv[64];
        double se[8];
        for (int64_t i = 0; i < de; i++) se[i] = SC(s_e, sc_off + i);
        for (int64_t i = 0; i < de; i++) damp[i] = (T)((lambda * A1(scaling_e, sc_off + i)) * (se[i] * se[i]));
        for (int64_t k = 0; k < de * de; k++) {
            int64_t i = k % de, j = k / de;
            Es[k] = FN(scaled)(A1(E_data, eo + k), se[i] * se[j]);
        }
        if (!FN(orc_inv_sym_block)(de, Es, damp, inv)) return 0;
        for (int64_t k = 0; k < de * de; k++) A1(E_inv, eo + k) = inv[k];
        /* z_e = E_inv * (b_e .* se): Float64 matvec, sum over j ascending */
        for (int64_t i = 0; i < de; i++) {
            double acc = 0.0;
            for (int64_t j = 0; j < de; j++) acc += (double)inv[j * de + i] * (A1(b_e, sc_off + j) * se[j]);
            A1(z, sc_off + i) = acc;
        }
        if (W) {
            for (int64_t c = A1(st->eliminated_ptr, e); c < A1(st->eliminated_ptr, e + 1); c++) {
                const int64_t co = A1(st->coffsets, c);
                const int64_t r = A1(st->entry_retained, c);
                const int64_t dr = A1(st->r_dims, r);
                const int64_t ro = A1(st->boffsets_r, r);
                for (int64_t ir = 1; ir <= dr; ir++) {
                    T col[8];
                    const double f = SC(s_r, ro + ir - 1);
                    for (int64_t i = 0; i < de; i++)
                        col[i] = FN(scaled)(A1(C_data, co + (ir - 1) * de + i), se[i] * f);
                    for (int64_t i = 0; i < de; i++) {
                        T acc = 0;
                        for (int64_t j = 0; j < de; j++) acc += inv[j * de + i] * col[j];
                        A1(W, co + (ir - 1) * de + i) = acc;
                    }
                }
            }
        }
        return 1;
    }
    /* dynamic path, de > 8 (src/schur_reduction.jl:164-225) */
    T *M = (T *)malloc(sizeof(T) * de * de);
    double *hii = (double *)malloc(sizeof(double) * de);
    for (int64_t j = 1; j <= de; j++)
        for (int64_t i = 1; i <= de; i++)
            M[(j - 1) * de + i - 1] = FN(scaled)(A1(E_data, eo + (j - 1) * de + i - 1),
                                                  SC(s_e, sc_off + i - 1) * SC(s_e, sc_off + j - 1));
    for (int64_t i = 1; i <= de; i++) {
        double si = SC(s_e, sc_off + i - 1);
        M[(i - 1) * de + i - 1] = (T)((double)M[(i - 1) * de + i - 1] + lambda * A1(scaling_e, sc_off + i - 1) * (si * si));
        hii[i - 1] = (double)A1(E_data, eo + (i - 1) * (de + 1)) * (si * si) + lambda * A1(scaling_e, sc_off + i - 1) * (si * si);
    }
    int ok = FN(chol_inverse_dyn)(de, M, hii);
    free(hii);
    if (!ok) { free(M); return 0; }
    for (int64_t k = 0; k < de * de; k++) A1(E_inv, eo + k) = M[k];
    free(M);
    for (int64_t i = 1; i <= de; i++) {
        double acc = 0.0;
        for (int64_t j = 1; j <= de; j++)
            acc += (double)A1(E_inv, eo + (j - 1) * de + i - 1) * (A1(b_e, sc_off + j - 1) * SC(s_e, sc_off + j - 1));
        A1(z, sc_off + i - 1) = acc;
    }
    if (W) {
        for (int64_t c = A1(st->eliminated_ptr, e); c < A1(st->eliminated_ptr, e + 1); c++) {
            const int64_t co = A1(st->coffsets, c);
            const int64_t r = A1(st->entry_retained, c);
            const int64_t dr = A1(st->r_dims, r);
            const int64_t ro = A1(st->boffsets_r, r);
            for (int64_t ir = 1; ir <= dr; ir++) {
                const int64_t row_off = co + (ir - 1) * de;
                const double f = SC(s_r, ro + ir - 1);
                for (int64_t i = 1; i <= de; i++) {
                    T acc = 0;
                    for (int64_t j = 1; j <= de; j++)
                        acc += A1(E_inv, eo + (j - 1) * de + i - 1) *
                               FN(scaled)(A1(C_data, row_off + j - 1), SC(s_e, sc_off + j - 1) * f);
                    A1(W, row_off + i - 1) = acc;
                }
            }
        }
    }
    return 1;
}

/* CAS-min of the lowest failing block (src/schur_reduction.jl:242-253) */
static void FN(record_failure)(int64_t *failed_block, int64_t e)
{
    int64_t old = __atomic_load_n(failed_block, __ATOMIC_SEQ_CST);
    while (old == 0 || e < old) {
        if (__atomic_compare_exchange_n(failed_block, &old, e, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) break;
    }
}

/* factor_all_eliminated! (src/schur_reduction.jl:439-455) over the flattened chunk ranges */
static int FN(factor_all)(const orc_schur *st, const T *E_data, const T *C_data, const double *b_e,
                          const double *scaling_e, double lambda, T *E_inv, double *z, T *W,
                          const double *s_r, const double *s_e, int64_t *failed_block, int nthreads)
{
    int64_t i0 = 0;
    while (i0 < st->nranges) {
        int64_t i1 = i0;
        while (i1 < st->nranges && st->chunk_run[i1] == st->chunk_run[i0]) i1++;
#pragma omp parallel for schedule(static, 1) num_threads(nthreads > 1 ? nthreads : 1) if (nthreads > 1)
        for (int64_t k = i0; k < i1; k++) {
            for (int64_t e = st->chunk_lo[k]; e <= st->chunk_hi[k]; e++) {
                int64_t fb = __atomic_load_n(failed_block, __ATOMIC_SEQ_CST);
                if (fb != 0 && e > fb) break;
                if (!FN(factor_eliminated)(st, E_inv, z, W, E_data, C_data, b_e, scaling_e, lambda, e, s_r, s_e)) {
                    FN(record_failure)(failed_block, e);
                    break;
                }
            }
        }
        if (*failed_block != 0) return 0;
        i0 = i1;
    }
    return 1;
}

/* dot_col of a scaled coupling column with z (Float64 accumulation, src/conditioning.jl:222-235) */
static inline double FN(dot_cz)(const T *C_data, int64_t off, int64_t de, const double *s_e, int64_t so, double f,
                                const double *z, int64_t zo)
{
    double acc = 0.0;
    for (int64_t i = 0; i < de; i++) acc += (double)FN(scaled)(A1(C_data, off + i), SC(s_e, so + i) * f) * A1(z, zo + i);
    return acc;
}

/* accumulate_entry_{dense,sparse}! (src/schur_reduction.jl:262-354) for coupling c2 of eliminated block e */
static void FN(accumulate_entry)(int sparse, const orc_schur *st, const orc_pattern *pat, T *S, double *g_S,
                                 const T *C_data, const T *W, const double *z, int64_t c2, int64_t e,
                                 const double *s_r, const double *s_e)
{
    const int64_t de = A1(st->e_dims, e);
    const int64_t r2 = A1(st->entry_retained, c2);
    const int64_t dr2 = A1(st->r_dims, r2);
    const int64_t ro2 = A1(st->boffsets_r, r2);
    const int64_t co2 = A1(st->coffsets, c2);
    const int64_t sc_off = A1(st->boffsets_e, e);
    for (int64_t ir = 1; ir <= dr2; ir++)
        A1(g_S, ro2 + ir - 1) -= FN(dot_cz)(C_data, co2 + (ir - 1) * de, de, s_e, sc_off, SC(s_r, ro2 + ir - 1), z, sc_off);
    int64_t col_base = 0, stride = 0, b_row_idx = 0, b_row_end = 0;
    if (sparse) {
        col_base = A1(pat->S_colptr, ro2);
        stride = A1(pat->col_heights, r2);
        b_row_idx = A1(pat->block_colptr, r2);
        b_row_end = A1(pat->block_colptr, r2 + 1) - 1;
    }
    for (int64_t c1 = A1(st->eliminated_ptr, e); c1 <= c2; c1++) {
        const int64_t r1 = A1(st->entry_retained, c1);
        const int64_t dr1 = A1(st->r_dims, r1);
        const int64_t wo1 = A1(st->coffsets, c1);
        int64_t row_offset = 0;
        if (sparse) {
            while (b_row_idx <= b_row_end && A1(pat->block_rowval, b_row_idx) < r1) b_row_idx++;
            if (!(b_row_idx <= b_row_end && A1(pat->block_rowval, b_row_idx) == r1)) abort(); /* throw_missing_block */
            row_offset = A1(pat->block_row_offsets, b_row_idx);
        }
        const int64_t ro1 = A1(st->boffsets_r, r1);
        for (int64_t col = 1; col <= dr2; col++) {
            const int64_t c_off = co2 + (col - 1) * de;
            const double f = SC(s_r, ro2 + col - 1);
            T cs[64];
            for (int64_t i = 0; i < de && i < 64; i++) cs[i] = FN(scaled)(A1(C_data, c_off + i), SC(s_e, sc_off + i) * f);
            const int64_t nrow = (r1 == r2) ? col : dr1;
            for (int64_t row = 1; row <= nrow; row++) {
                T acc = 0;
                for (int64_t i = 0; i < de; i++)
                    acc += A1(W, wo1 + (row - 1) * de + i) *
                           (de <= 64 ? cs[i] : FN(scaled)(A1(C_data, c_off + i), SC(s_e, sc_off + i) * f));
                if (sparse) {
                    const int64_t col_ptr = col_base + (col - 1) * stride + row_offset;
                    A1(S, col_ptr + row - 1) -= acc;
                } else {
                    S[(ro2 + col - 2) * st->nr + (ro1 + row - 2)] -= acc;
                }
            }
        }
    }
}

/* reduce_schur_dense! / reduce_schur_sparse! (src/schur_reduction.jl:465-507); validation is the caller's.
 * W may be NULL only if the caller does not need it -- the reduction itself needs it, so it must be given. */
int FN(orc_reduce_schur)(int sparse, const orc_schur *st, const orc_pattern *pat, T *S, double *g_S,
                         const double *b_r, const double *b_e, const double *scaling_e, double lambda,
                         const T *E_data, const T *C_data, T *E_inv, double *z, T *W,
                         const double *s_r, const double *s_e, int64_t *failed_block, int nthreads)
{
    *failed_block = 0;
    if (!FN(factor_all)(st, E_data, C_data, b_e, scaling_e, lambda, E_inv, z, W, s_r, s_e, failed_block, nthreads))
        return 0;
    for (int64_t i = 1; i <= st->nr; i++) A1(g_S, i) = A1(b_r, i) * SC(s_r, i); /* scaled_copy! */
#pragma omp parallel for schedule(static, 1) num_threads(nthreads > 1 ? nthreads : 1) if (nthreads > 1)
    for (int64_t k = 0; k < st->ncol_chunks; k++) {
        for (int64_t r2 = st->col_lo[k]; r2 <= st->col_hi[k]; r2++) {
            for (int64_t kk = A1(st->retained_ptr, r2); kk < A1(st->retained_ptr, r2 + 1); kk++) {
                const int64_t c2 = A1(st->retained_entry, kk);
                const int64_t e = A1(st->entry_eliminated, c2);
                FN(accumulate_entry)(sparse, st, pat, S, g_S, C_data, W, z, c2, e, s_r, s_e);
            }
        }
    }
    return 1;
}

/* reset_schur_dense! (src/schur_reduction.jl:389-407) */
int FN(orc_reset_schur_dense)(int64_t nr, T *S, const T *R, const double *scaling_r, double lambda, const double *s_r)
{
    for (int64_t j = 1; j <= nr; j++)
        for (int64_t i = 1; i <= nr; i++) {
            const int64_t k = (j - 1) * nr + (i - 1);
            S[k] = s_r ? FN(scaled)(R[k], A1(s_r, i) * A1(s_r, j)) : R[k];
        }
    for (int64_t i = 1; i <= nr; i++) {
        const double si = SC(s_r, i);
        const int64_t k = (i - 1) * nr + (i - 1);
        S[k] = (T)((double)S[k] + lambda * A1(scaling_r, i) * (si * si));
    }
    return 0;
}

/* reset_schur_sparse! (src/schur_reduction.jl:416-436) */
int FN(orc_reset_schur_sparse)(int64_t nnz, T *nzval, int64_t nR, const T *R_nzval, const int64_t *r_to_s_map,
                               int64_t nr, const int64_t *diag_inds, const double *scaling_r, double lambda,
                               const double *s_r, const int64_t *S_colptr, const int64_t *S_rowval)
{
    for (int64_t i = 0; i < nnz; i++) nzval[i] = 0;
    for (int64_t i = 1; i <= nR; i++) {
        const int64_t idx = A1(r_to_s_map, i);
        if (idx > 0) A1(nzval, idx) = A1(R_nzval, i);
    }
    if (s_r) { /* scale_sparse! (src/conditioning.jl:247-253) */
        for (int64_t j = 1; j <= nr; j++)
            for (int64_t k = A1(S_colptr, j); k < A1(S_colptr, j + 1); k++)
                A1(nzval, k) = FN(scaled)(A1(nzval, k), A1(s_r, A1(S_rowval, k)) * A1(s_r, j));
    }
    for (int64_t i = 1; i <= nr; i++) {
        const int64_t idx = A1(diag_inds, i);
        const double si = SC(s_r, i);
        A1(nzval, idx) = (T)((double)A1(nzval, idx) + lambda * A1(scaling_r, i) * (si * si));
    }
    return 0;
}

/* backsubstitute! (src/schur_reduction.jl:519-614): x_e = z_e - E_inv * sum_r C_er x_r (t in Float64) */
void FN(orc_backsubstitute_schur)(const orc_schur *st, double *x_e, const double *x_r, const T *C_data,
                                  const T *E_inv, const double *z, const double *s_r, const double *s_e)
{
    double t[4096];
    for (int64_t e = 1; e <= st->neblocks; e++) {
        const int64_t de = A1(st->e_dims, e);
        const int64_t e_off = A1(st->boffsets_e, e);
        const int64_t eo = A1(st->eoffsets, e);
        for (int64_t i = 0; i < de; i++) t[i] = 0.0;
        for (int64_t c = A1(st->eliminated_ptr, e); c < A1(st->eliminated_ptr, e + 1); c++) {
            const int64_t r = A1(st->entry_retained, c);
            const int64_t dr = A1(st->r_dims, r);
            const int64_t ro = A1(st->boffsets_r, r);
            const int64_t co = A1(st->coffsets, c);
            for (int64_t ir = 1; ir <= dr; ir++) {
                const double xc = A1(x_r, ro + ir - 1);
                const double f = SC(s_r, ro + ir - 1);
                for (int64_t i = 0; i < de; i++)
                    t[i] += (double)FN(scaled)(A1(C_data, co + (ir - 1) * de + i), SC(s_e, e_off + i) * f) * xc;
            }
        }
        for (int64_t i = 0; i < de; i++) {
            if (de > 8) { /* run-time-size path (src/schur_reduction.jl:578-614): z - each term in turn */
                double acc = A1(z, e_off + i);
                for (int64_t j = 0; j < de; j++) acc -= (double)A1(E_inv, eo + j * de + i) * t[j];
                A1(x_e, e_off + i) = acc;
            } else {
                double acc = 0.0; /* E_inv * t, then z - that (static path: SMatrix * SVector) */
                for (int64_t j = 0; j < de; j++) acc += (double)A1(E_inv, eo + j * de + i) * t[j];
                A1(x_e, e_off + i) = A1(z, e_off + i) - acc;
            }
        }
    }
}

/* copyto!(dest, src, scatter_map) (src/SparseBlockLinearSystem.jl:356-362): a gather */
void FN(orc_copyto)(int64_t n, T *dest, const T *src, const int64_t *scatter_map)
{
    for (int64_t i = 1; i <= n; i++) A1(dest, i) = A1(src, A1(scatter_map, i));
}

/* ---------------------------------------------------------------------------------------------------------
 * QR elimination (src/qr_reduction.jl)
 * ------------------------------------------------------------------------------------------------------- */

/* copy_table! (src/qr_reduction.jl:192-255); cs (scale over all variables) NULL = UnitScale */
static int FN(copy_table)(const orc_schur *st, const orc_jacobian *jst, T *F, double *Fr, const T *Rw, const double *Rr,
                          int64_t t, const double *cs)
{
    const int64_t M = A1(jst->table_nrows, t);
    const int64_t C = A1(jst->table_ncols, t) - 1;
    const int64_t off = A1(jst->raw_offsets, t);
    T bad = 0;
    if (!cs) {
        for (int64_t i = off; i <= off + M * C - 1; i++) {
            T x = A1(Rw, i);
            A1(F, i) = x;
            bad += x - x;
        }
    } else {
        for (int64_t i = off; i <= off + M * C - 1; i++) { T x = A1(Rw, i); bad += x - x; }
        for (int64_t u = A1(jst->table_ptr, t); u < A1(jst->table_ptr, t + 1); u++) {
            const int64_t r = A1(jst->entry_retained, u);
            const int64_t ro = A1(st->boffsets_r, r);
            for (int64_t k = 1; k <= A1(st->r_dims, r); k++) {
                const double f = A1(cs, ro + k - 1);
                const int64_t cb = off + (A1(jst->entry_col, u) + k - 2) * M;
                for (int64_t i = cb; i <= cb + M - 1; i++) A1(F, i) = FN(scaled)(A1(Rw, i), f);
            }
        }
        if (t <= st->neblocks) {
            const int64_t de = A1(st->e_dims, t);
            const int64_t eo = st->nr + A1(st->boffsets_e, t) - 1;
            for (int64_t k = 1; k <= de; k++) {
                const double f = A1(cs, eo + k);
                const int64_t cb = off + (C - de + k - 1) * M;
                for (int64_t i = cb; i <= cb + M - 1; i++) A1(F, i) = FN(scaled)(A1(Rw, i), f);
            }
        }
    }
    const int64_t rb = A1(jst->residual_offsets, t);
    double badr = 0.0;
    for (int64_t i = rb; i <= rb + M - 1; i++) {
        double x = A1(Rr, i);
        A1(Fr, i) = x;
        badr += x - x;
    }
    return !isnan(bad) && !isnan(badr);
}

/* qr_table! (src/qr_reduction.jl:50-131) */
static int FN(qr_table)(const orc_schur *st, const orc_jacobian *jst, T *F, double *Fr, const T *Rw, const double *Rr,
                        const double *scaling_e, double lambda, int64_t e, const double *cs)
{
    const int64_t de = A1(st->e_dims, e);
    const int64_t M = A1(jst->table_nrows, e);
    const int64_t C = A1(jst->table_ncols, e) - 1;
    const int64_t off = A1(jst->raw_offsets, e);
    const int64_t rq = A1(jst->residual_offsets, e) - 1;
    const int64_t me = M - de;
    const int64_t ce = C - de;
    if (!FN(copy_table)(st, jst, F, Fr, Rw, Rr, e, cs)) return 0;
    for (int64_t j = 1; j <= C; j++)
        for (int64_t i = me + 1; i <= M; i++) A1(F, off + (j - 1) * M + i - 1) = 0;
    for (int64_t i = me + 1; i <= M; i++) A1(Fr, rq + i) = 0.0;
    const int64_t so = A1(st->boffsets_e, e);
    const int64_t nr = st->nr;
    for (int64_t k = 1; k <= de; k++)
        A1(F, off + (ce + k - 1) * M + me + k - 1) = (T)(sqrt(lambda * A1(scaling_e, so + k - 1)) * SC(cs, nr + so + k - 1));
    const T eps = sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07;
    const T tol = (T)M * eps;
    for (int64_t k = 1; k <= de; k++) {
        const int64_t cb = off + (ce + k - 1) * M - 1;
        T s_top = 0;
        for (int64_t i = 1; i <= k - 1; i++) s_top += A1(F, cb + i) * A1(F, cb + i);
        T s_sub = 0;
        for (int64_t i = k; i <= M; i++) s_sub += A1(F, cb + i) * A1(F, cb + i);
        T nrm = (T)sqrt((double)s_sub);
        T full = (T)sqrt((double)(s_top + s_sub));
        if (sizeof(T) == 4) { nrm = sqrtf((float)s_sub); full = sqrtf((float)(s_top + s_sub)); }
        T a = A1(F, cb + k);
        T beta = a >= 0 ? -nrm : nrm;
        if (!(isfinite(beta) && fabs((double)beta) > (double)(tol * full))) return 0;
        T scale = (T)1 / (a - beta);
        for (int64_t i = k + 1; i <= M; i++) A1(F, cb + i) *= scale;
        T tau = (beta - a) / beta;
        A1(F, cb + k) = beta;
        for (int64_t j = 1; j <= C; j++) {
            if (ce + 1 <= j && j <= ce + k) continue;
            const int64_t yb = off + (j - 1) * M - 1;
            T s = A1(F, yb + k);
            for (int64_t i = k + 1; i <= M; i++) s += A1(F, cb + i) * A1(F, yb + i);
            s *= tau;
            A1(F, yb + k) -= s;
            for (int64_t i = k + 1; i <= M; i++) A1(F, yb + i) -= s * A1(F, cb + i);
        }
        double s = A1(Fr, rq + k);
        for (int64_t i = k + 1; i <= M; i++) s += (double)A1(F, cb + i) * A1(Fr, rq + i);
        s *= (double)tau;
        A1(Fr, rq + k) -= s;
        for (int64_t i = k + 1; i <= M; i++) A1(Fr, rq + i) -= s * (double)A1(F, cb + i);
        for (int64_t i = k + 1; i <= M; i++) A1(F, cb + i) = 0;
    }
    return 1;
}

/* reduce_qr! (src/qr_reduction.jl:162-187); validation is the caller's */
int FN(orc_reduce_qr)(const orc_schur *st, const orc_jacobian *jst, T *F, double *Fr, const T *Rw, const double *Rr,
                      const double *scaling_e, double lambda, const double *cs, int64_t *failed_block, int nthreads)
{
    *failed_block = 0;
    int64_t i0 = 0;
    while (i0 < st->nranges) {
        int64_t i1 = i0;
        while (i1 < st->nranges && st->chunk_run[i1] == st->chunk_run[i0]) i1++;
#pragma omp parallel for schedule(static, 1) num_threads(nthreads > 1 ? nthreads : 1) if (nthreads > 1)
        for (int64_t k = i0; k < i1; k++) {
            for (int64_t e = st->chunk_lo[k]; e <= st->chunk_hi[k]; e++) {
                int64_t fb = __atomic_load_n(failed_block, __ATOMIC_SEQ_CST);
                if (fb != 0 && e > fb) break;
                if (!FN(qr_table)(st, jst, F, Fr, Rw, Rr, scaling_e, lambda, e, cs)) {
                    FN(record_failure)(failed_block, e);
                    break;
                }
            }
        }
        if (*failed_block != 0) return 0;
        i0 = i1;
    }
    for (int64_t t = st->neblocks + 1; t <= jst->ntables; t++) FN(copy_table)(st, jst, F, Fr, Rw, Rr, t, cs);
    return 1;
}

/* qr_backsubstitute_group! (src/qr_reduced.jl:24-95): x_e = U^-1 (q - B x_r), accumulated in double */
void FN(orc_qr_backsubstitute)(const orc_schur *st, const orc_jacobian *jst, double *x_e, const double *x_r,
                               const T *F, const double *Fr)
{
    double t[4096];
    for (int64_t e = 1; e <= st->neblocks; e++) {
        const int64_t de = A1(st->e_dims, e);
        const int64_t M = A1(jst->table_nrows, e);
        const int64_t C = A1(jst->table_ncols, e);
        const int64_t off = A1(jst->raw_offsets, e);
        const int64_t ce = C - de - 1;
        const int64_t qb = A1(jst->residual_offsets, e);
        for (int64_t i = 0; i < de; i++) t[i] = A1(Fr, qb + i);
        for (int64_t u = A1(jst->table_ptr, e); u < A1(jst->table_ptr, e + 1); u++) {
            const int64_t r = A1(jst->entry_retained, u);
            const int64_t ro = A1(st->boffsets_r, r);
            const int64_t cb = off + (A1(jst->entry_col, u) - 1) * M;
            for (int64_t j = 1; j <= A1(st->r_dims, r); j++) {
                const double xj = A1(x_r, ro + j - 1);
                for (int64_t i = 0; i < de; i++) t[i] -= (double)A1(F, cb + (j - 1) * M + i) * xj;
            }
        }
        const int64_t eo = A1(st->boffsets_e, e);
        for (int64_t i = de; i >= 1; i--) {
            double acc = t[i - 1];
            for (int64_t j = i + 1; j <= de; j++) acc -= (double)A1(F, off + (ce + j - 1) * M + i - 1) * A1(x_e, eo + j - 1);
            A1(x_e, eo + i - 1) = acc / (double)A1(F, off + (ce + i - 1) * M + i - 1);
        }
    }
}

static inline double FN(dot_run)(const T *a, int64_t ia, const T *b, int64_t ib, int64_t n)
{
    T acc = 0;
    for (int64_t i = 0; i < n; i++) acc += A1(a, ia + i) * A1(b, ib + i);
    return (double)acc;
}
static inline double FN(dot_run_r)(const T *a, int64_t ia, const double *b, int64_t ib, int64_t n)
{
    double acc = 0;
    for (int64_t i = 0; i < n; i++) acc += (double)A1(a, ia + i) * A1(b, ib + i);
    return acc;
}

/* reduced_normal_dense! / reduced_normal_sparse! (src/qr_reduced.jl:373-480): upper of S = J~'J~, g_S = J~'r~ */
void FN(orc_reduced_normal)(int sparse, const orc_schur *st, const orc_jacobian *jst, const orc_pattern *pat,
                            T *S, double *g_S, const T *F, const double *Fr)
{
    const int64_t nr = st->nr;
    for (int64_t r2 = 1; r2 <= st->nrblocks; r2++) {
        const int64_t ro2 = A1(st->boffsets_r, r2);
        const int64_t dr2 = A1(st->r_dims, r2);
        if (sparse) {
            for (int64_t i = A1(pat->S_colptr, ro2); i < A1(pat->S_colptr, ro2 + dr2); i++) A1(S, i) = 0;
        } else {
            for (int64_t col = ro2; col <= ro2 + dr2 - 1; col++)
                for (int64_t row = 1; row <= nr; row++) S[(col - 1) * nr + row - 1] = 0;
        }
        for (int64_t col = ro2; col <= ro2 + dr2 - 1; col++) A1(g_S, col) = 0.0;
        for (int64_t k = A1(jst->retained_ptr, r2); k < A1(jst->retained_ptr, r2 + 1); k++) {
            const int64_t u2 = A1(jst->retained_entry, k);
            const int64_t t = A1(jst->entry_table, u2);
            const int64_t M = A1(jst->table_nrows, t);
            const int64_t top = A1(jst->table_top, t);
            const int64_t n = M - top;
            const int64_t off = A1(jst->raw_offsets, t) + top;
            const int64_t cb2 = off + (A1(jst->entry_col, u2) - 1) * M;
            const int64_t rb = A1(jst->residual_offsets, t) + top;
            for (int64_t j = 1; j <= dr2; j++) A1(g_S, ro2 + j - 1) += FN(dot_run_r)(F, cb2 + (j - 1) * M, Fr, rb, n);
            for (int64_t u1 = A1(jst->table_ptr, t); u1 <= u2; u1++) {
                const int64_t r1 = A1(jst->entry_retained, u1);
                const int64_t dr1 = A1(st->r_dims, r1);
                const int64_t cb1 = off + (A1(jst->entry_col, u1) - 1) * M;
                int64_t row_offset = 0;
                if (sparse) {
                    int64_t lo = A1(pat->block_colptr, r2), hi = A1(pat->block_colptr, r2 + 1) - 1, idx = lo;
                    while (idx <= hi && A1(pat->block_rowval, idx) < r1) idx++;
                    if (!(idx <= hi && A1(pat->block_rowval, idx) == r1)) abort();
                    row_offset = A1(pat->block_row_offsets, idx);
                }
                for (int64_t col = 1; col <= dr2; col++) {
                    const int64_t nrow = (u1 == u2) ? col : dr1;
                    for (int64_t row = 1; row <= nrow; row++) {
                        T v = (T)FN(dot_run)(F, cb1 + (row - 1) * M, F, cb2 + (col - 1) * M, n);
                        if (sparse)
                            A1(S, A1(pat->S_colptr, ro2) + (col - 1) * A1(pat->col_heights, r2) + row_offset + row - 1) += v;
                        else
                            S[(ro2 + col - 2) * nr + A1(st->boffsets_r, r1) + row - 2] += v;
                    }
                }
            }
        }
    }
}

/* reduced_diagonal_blocks! (src/qr_reduced.jl:484-512) */
void FN(orc_reduced_diagonal_blocks)(const orc_schur *st, const orc_jacobian *jst, T *D, const T *F)
{
    for (int64_t r = 1; r <= st->nrblocks; r++) {
        const int64_t dr = A1(st->r_dims, r);
        const int64_t d0 = A1(jst->doffsets, r);
        for (int64_t i = d0; i <= d0 + dr * dr - 1; i++) A1(D, i) = 0;
        for (int64_t k = A1(jst->retained_ptr, r); k < A1(jst->retained_ptr, r + 1); k++) {
            const int64_t u = A1(jst->retained_entry, k);
            const int64_t t = A1(jst->entry_table, u);
            const int64_t M = A1(jst->table_nrows, t);
            const int64_t top = A1(jst->table_top, t);
            const int64_t cb = A1(jst->raw_offsets, t) + top + (A1(jst->entry_col, u) - 1) * M;
            for (int64_t col = 1; col <= dr; col++)
                for (int64_t row = 1; row <= col; row++)
                    A1(D, d0 + (col - 1) * dr + row - 1) += (T)FN(dot_run)(F, cb + (row - 1) * M, F, cb + (col - 1) * M, M - top);
        }
        for (int64_t col = 1; col <= dr; col++)
            for (int64_t row = col + 1; row <= dr; row++)
                A1(D, d0 + (col - 1) * dr + row - 1) = A1(D, d0 + (row - 1) * dr + col - 1);
    }
}

/* reduced_mul! (src/qr_reduced.jl:161-193): y = J~ v, accumulated in y's type (double) */
void FN(orc_reduced_mul)(const orc_schur *st, const orc_jacobian *jst, double *y, const T *F, const double *v)
{
    for (int64_t t = 1; t <= jst->ntables; t++) {
        const int64_t M = A1(jst->table_nrows, t);
        const int64_t top = A1(jst->table_top, t);
        const int64_t off = A1(jst->raw_offsets, t);
        const int64_t yo = A1(jst->reduced_offsets, t);
        const int64_t n = M - top;
        for (int64_t i = 0; i < n; i++) A1(y, yo + i) = 0.0;
        for (int64_t u = A1(jst->table_ptr, t); u < A1(jst->table_ptr, t + 1); u++) {
            const int64_t r = A1(jst->entry_retained, u);
            const int64_t ro = A1(st->boffsets_r, r);
            const int64_t cb = off + (A1(jst->entry_col, u) - 1) * M + top;
            for (int64_t j = 1; j <= A1(st->r_dims, r); j++) {
                const double vj = A1(v, ro + j - 1);
                const int64_t c0 = cb + (j - 1) * M;
                for (int64_t i = 0; i < n; i++) A1(y, yo + i) += (double)A1(F, c0 + i) * vj;
            }
        }
    }
}

/* reduced_tmul! (src/qr_reduced.jl:195-223): g = J~' w */
void FN(orc_reduced_tmul)(const orc_schur *st, const orc_jacobian *jst, double *g, const T *F, const double *w)
{
    for (int64_t r = 1; r <= st->nrblocks; r++) {
        const int64_t ro = A1(st->boffsets_r, r);
        const int64_t dr = A1(st->r_dims, r);
        for (int64_t j = 0; j < dr; j++) A1(g, ro + j) = 0.0;
        for (int64_t k = A1(jst->retained_ptr, r); k < A1(jst->retained_ptr, r + 1); k++) {
            const int64_t u = A1(jst->retained_entry, k);
            const int64_t t = A1(jst->entry_table, u);
            const int64_t M = A1(jst->table_nrows, t);
            const int64_t top = A1(jst->table_top, t);
            const int64_t cb = A1(jst->raw_offsets, t) + (A1(jst->entry_col, u) - 1) * M + top;
            const int64_t wo = A1(jst->reduced_offsets, t);
            for (int64_t j = 1; j <= dr; j++) A1(g, ro + j - 1) += FN(dot_run_r)(F, cb + (j - 1) * M, w, wo, M - top);
        }
    }
}

/* reduced_residual! (src/qr_reduced.jl:227-243) */
void FN(orc_reduced_residual)(const orc_jacobian *jst, double *rt, const double *Fr)
{
    for (int64_t t = 1; t <= jst->ntables; t++) {
        const int64_t M = A1(jst->table_nrows, t);
        const int64_t top = A1(jst->table_top, t);
        const int64_t cb = A1(jst->residual_offsets, t) + top;
        const int64_t ro = A1(jst->reduced_offsets, t);
        for (int64_t i = 0; i <= M - top - 1; i++) A1(rt, ro + i) = A1(Fr, cb + i);
    }
}

/* factor_diagonal_blocks! (src/diagonal_blocks.jl:6-30): lower Cholesky in place; returns 1 on success */
int FN(orc_factor_diagonal_blocks)(int64_t nblocks, const int64_t *dims, T *D)
{
    int64_t o = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t d = dims[b];
        for (int64_t j = 1; j <= d; j++) {
            T s = D[o + (j - 1) * d + j - 1];
            for (int64_t k = 1; k <= j - 1; k++) { T l = D[o + (k - 1) * d + j - 1]; s -= l * l; }
            if (!(s > 0 && isfinite(s))) return 0;
            T ljj = (T)sqrt((double)s);
            if (sizeof(T) == 4) ljj = sqrtf((float)s);
            D[o + (j - 1) * d + j - 1] = ljj;
            for (int64_t i = j + 1; i <= d; i++) {
                T t = D[o + (j - 1) * d + i - 1];
                for (int64_t k = 1; k <= j - 1; k++) t -= D[o + (k - 1) * d + i - 1] * D[o + (k - 1) * d + j - 1];
                D[o + (j - 1) * d + i - 1] = t / ljj;
            }
        }
        o += d * d;
    }
    return 1;
}

/* solve_diagonal_blocks! (src/diagonal_blocks.jl:33-61): y = (L L')^-1 x block by block */
void FN(orc_solve_diagonal_blocks)(int64_t nblocks, const int64_t *dims, double *y, const T *D, const double *x)
{
    int64_t o = 0, p = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t d = dims[b];
        if (y != x)
            for (int64_t i = 1; i <= d; i++) y[p + i - 1] = x[p + i - 1];
        for (int64_t j = 1; j <= d; j++) {
            double yj = y[p + j - 1] / (double)D[o + (j - 1) * d + j - 1];
            y[p + j - 1] = yj;
            for (int64_t i = j + 1; i <= d; i++) y[p + i - 1] -= (double)D[o + (j - 1) * d + i - 1] * yj;
        }
        for (int64_t j = d; j >= 1; j--) {
            double t = y[p + j - 1];
            for (int64_t i = j + 1; i <= d; i++) t -= (double)D[o + (j - 1) * d + i - 1] * y[p + i - 1];
            y[p + j - 1] = t / (double)D[o + (j - 1) * d + j - 1];
        }
        o += d * d;
        p += d;
    }
}

#undef FN
