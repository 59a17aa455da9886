// QR elimination on gfx950: reduce_qr! (src/qr_reduction.jl:162-187) behind the C ABI of include/bls_amd.h.
//
// Each table T_e = [J_r | J_e] (M = m_e + d_e rows, column-major; residual in its own Float64 vector) is worked
// by a row of 16 lanes, lane l holding rows l, l + 16, ...: the lanes compute the d_e Householder reflectors of
// qr_table! (src/qr_reduction.jl:50-131) -- norms and dot products by in-lane partial sums and DPP row sums, the
// singularity test |beta| > M eps ||col|| -- and apply them to every retained column and the residual, each read
// and written as coalesced runs of 16 values. Tables are bucketed by row count (at most 16, 32 or 64 rows: 1, 2 or
// 4 rows per lane, four tables per wavefront) and sorted widest first within a bucket. The raw tables are never
// modified (retry semantics of the reference).
#include <algorithm>
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {

constexpr int QR_MAXDE = 8;        // largest d_e of the register kernels
constexpr int QRW_MAXDE = 4096;    // largest d_e of the workgroup kernels (backsubstitute's t in LDS)

// T: element type of the raw and factored tables (SchurBlockJacobian{T}); the residual is Float64 and the
// reflectors are formed and applied in Float64 registers for either T, with the reference's singularity test
// |beta| > M eps(T) ||col||.
template <class T> struct QRArgs {
    const int32_t *tables;    // tables of this bucket (0-based)
    int64_t ntab;
    const int32_t *nrows, *ncols, *e_dims;   // per table: M, raw columns (table_ncols - 1); e_dims per table (0 = free)
    const int64_t *raw_off, *res_off, *boff_e;
    const T *Rw;
    const double *Rr, *scaling_e;
    double lambda;
    T *F;
    double *Fr;
    unsigned long long *flag;
    // conditioning (scale_table_columns!, src/qr_reduction.jl:236-255): cs over all nr + ne variables, NULL =
    // UnitScale; a retained column finds its variable through the table's entries
    const double *cs;
    int64_t nr;
    const int32_t *table_ptr, *entry_retained, *entry_col, *r_dims;
    const int64_t *boff_r;
};

// The conditioning factor of retained column j (0-based) of table t: s of the variable it holds.
template <class T> __device__ inline double column_scale(const QRArgs<T> &a, int t, int j) {
    for (int u = a.table_ptr[t]; u < a.table_ptr[t + 1]; u++) {
        const int c0 = a.entry_col[u], r = a.entry_retained[u];
        if (j >= c0 && j < c0 + a.r_dims[r]) return a.cs[a.boff_r[r] + (j - c0)];
    }
    return 1.0;   // not reached for a column of the structure
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// A 64-bit DPP move within a row of 16 lanes (two 32-bit moves); every lane of the row is active where it is used.
template <int CTRL> __device__ inline double dpp_mov(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a row, returned to all of them: partners at xor 1 and xor 2 (quad_perm), then the other
// half of each 8 (row_half_mirror) and the other 8 (row_mirror). Each step adds the same two partials in every
// lane, so all 16 lanes hold the same bits.
__device__ inline double row_sum16(double v) {
    v += dpp_mov<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_mov<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_mov<0x141>(v);   // row_half_mirror
    v += dpp_mov<0x140>(v);   // row_mirror
    return v;
}

// One table per row of 16 lanes, four per wavefront: lane l of a table's row holds rows l, l + 16, ... (R of them;
// bucket R takes tables of at most 16 R rows), so a column is R coalesced runs of 16 values. The lanes first form the
// d_e reflectors from the eliminated columns (norms and dot products by row sums, the singularity test
// |beta| > M eps(T) ||col||), keeping v_k and tau_k in registers; they then stream the retained columns and the
// residual NC at a time, applying H_1 ... H_de to each in the reference's sequence. DEM bounds d_e (3 or 8).
template <int R, int DEM, class T>
__global__ __launch_bounds__(256) void k_qr_tables(QRArgs<T> a) {
    constexpr int NC = R == 4 ? 2 : 4;   // columns a lane holds at once
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int l = lane & 15;
    const int base = lane & ~15;
    const int64_t first = ((int64_t)blockIdx.x * 4 + w) * 4;
    if (first >= a.ntab) return;   // the whole wavefront is past the last table
    const int64_t ti = first + (lane >> 4);
    if (ti >= a.ntab) return;      // this row of lanes has no table
    const int t = a.tables[ti];
    const int M = a.nrows[t];
    const int C = a.ncols[t];
    const int de = a.e_dims[t];
    const int64_t off = a.raw_off[t];
    const int64_t rq = a.res_off[t];
    if (de == 0) {   // a free table: copied (copy_table!, its result unused by reduce_qr!), columns scaled
        for (int j = 0; j < C; j++) {
            const double f = a.cs ? column_scale(a, t, j) : 1.0;
            for (int i = l; i < M; i += 16) {
                const T x = a.Rw[off + (int64_t)j * M + i];
                a.F[off + (int64_t)j * M + i] = a.cs ? (T)((double)x * f) : x;
            }
        }
        for (int i = l; i < M; i += 16) a.Fr[rq + i] = a.Rr[rq + i];
        return;
    }
    const int me = M - de, ce = C - de;
    const int64_t so = a.boff_e[t];
    double bad = 0.0;
    // panel: rows l + 16 r of the eliminated columns (damping rows zero, sqrt(lambda sc) s_e on their diagonal)
    double P[DEM][R];
#pragma unroll
    for (int k = 0; k < DEM; k++) {
        const double f = (k < de && a.cs) ? a.cs[a.nr + so + k] : 1.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = r * 16 + l;
            double v = 0.0;
            if (k < de) {
                if (row < me) {
                    const double x = (double)a.Rw[off + (int64_t)(ce + k) * M + row];
                    v = (double)(T)(x * f);
                    bad += x - x;
                } else if (row == me + k) {
                    v = (double)(T)(sqrt(a.lambda * a.scaling_e[so + k]) * f);
                }
            }
            P[k][r] = v;
        }
    }
    const double tol = (double)M * (sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-07);
    double V[DEM][R], tau[DEM];
    bool singular = false;
#pragma unroll
    for (int k = 0; k < DEM; k++) {
#pragma unroll
        for (int r = 0; r < R; r++) V[k][r] = 0.0;
        tau[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < DEM; k++) {
        if (k >= de) break;
        double top = 0.0, sub = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = r * 16 + l;
            const double x = P[k][r];
            if (row < k) top += x * x;
            else sub += x * x;
        }
        const double s_top = row_sum16(top), s_sub = row_sum16(sub);
        const double nrm = sqrt(s_sub);
        const double full = sqrt(s_top + s_sub);
        const double akk = __shfl(P[k][0], base + k, 64);   // row k (< 16) sits in lane k, slot 0
        const double beta = akk >= 0.0 ? -nrm : nrm;
        if (!(isfinite(beta) && fabs(beta) > tol * full)) { singular = true; break; }
        const double scale = 1.0 / (akk - beta);
        const double tk = (beta - akk) / beta;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = r * 16 + l;
            const double x = P[k][r];
            V[k][r] = row > k ? x * scale : (row == k ? 1.0 : 0.0);
            P[k][r] = row == k ? beta : (row < k ? x : 0.0);
        }
        tau[k] = tk;
        // apply H_k to the eliminated columns not yet triangularised
#pragma unroll
        for (int j = 0; j < DEM; j++) {
            if (j > k && j < de) {
                double part = 0.0;
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (r * 16 + l > k) part += V[k][r] * P[j][r];
                const double yk = __shfl(P[j][0], base + k, 64);
                const double sj = (row_sum16(part) + yk) * tk;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int row = r * 16 + l;
                    if (row == k) P[j][r] -= sj;
                    else if (row > k) P[j][r] -= sj * V[k][r];
                }
            }
        }
    }
    const unsigned long long gmask = 0xFFFFull << base;
    if (singular || (__ballot(isnan(bad)) & gmask)) {
        if (l == 0) atomicMin(a.flag, (unsigned long long)(t + 1));
        return;
    }
    // eliminated columns of F: [U; 0]
    for (int k = 0; k < de; k++)
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = r * 16 + l;
            if (row < M) a.F[off + (int64_t)(ce + k) * M + row] = (T)P[k][r];
        }
    // retained columns and the residual (column ce), NC at a time
    for (int j0 = 0; j0 <= ce; j0 += NC) {
        double x[NC][R];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int j = j0 + c;
            const double f = (a.cs && j < ce) ? column_scale(a, t, j) : 1.0;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int row = r * 16 + l;
                double xv = 0.0;
                if (j < ce) {
                    if (row < me) xv = (double)a.Rw[off + (int64_t)j * M + row];
                    bad += xv - xv;
                    if (a.cs) xv = (double)(T)(xv * f);
                } else if (j == ce) {
                    if (row < me) xv = a.Rr[rq + row];
                    bad += xv - xv;
                }
                x[c][r] = xv;
            }
        }
#pragma unroll
        for (int k = 0; k < DEM; k++) {
            if (k >= de) break;
            double sc[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) {
                double p = 0.0;
#pragma unroll
                for (int r = 0; r < R; r++) p += V[k][r] * x[c][r];
                sc[c] = row_sum16(p);
            }
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int r = 0; r < R; r++) x[c][r] -= (tau[k] * sc[c]) * V[k][r];
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int j = j0 + c;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int row = r * 16 + l;
                if (row < M) {
                    if (j < ce) a.F[off + (int64_t)j * M + row] = (T)x[c][r];
                    else if (j == ce) a.Fr[rq + row] = x[c][r];
                }
            }
        }
    }
    if ((__ballot(isnan(bad)) & gmask) && l == 0) atomicMin(a.flag, (unsigned long long)(t + 1));
}

// Tables the register kernel does not take (more than 64 rows, or d_e > QR_MAXDE): one workgroup per table, the
// table and its residual staged in a Float64 workspace (column-major, M x (C + 1)), the reflectors of qr_table!
// (src/qr_reduction.jl:50-131) formed by workgroup reductions (a fixed butterfly per wavefront, then the four
// wavefronts in order) and applied one column per wavefront. Same values, test and singularity rule as k_qr_tables.
constexpr int QRW_THREADS = 256;
template <class T>
__global__ __launch_bounds__(QRW_THREADS) void k_qr_table_wide(QRArgs<T> a, double *ws, int64_t wsz) {
    __shared__ double red[2][QRW_THREADS / 64];
    __shared__ int bad_sh;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    double *X = ws + (int64_t)blockIdx.x * wsz;
    for (int64_t ti = blockIdx.x; ti < a.ntab; ti += gridDim.x) {
        const int t = a.tables[ti];
        const int M = a.nrows[t], C = a.ncols[t], de = a.e_dims[t];
        const int64_t off = a.raw_off[t], rq = a.res_off[t], so = a.boff_e[t];
        const int me = M - de, ce = C - de;
        if (tid == 0) bad_sh = 0;
        __syncthreads();
        double bad = 0.0;
        // stage: retained columns (scaled), eliminated columns with the damping rows sqrt(lambda sc) s_e on their
        // diagonal, the residual; raw rows below me are the damping rows and read as zero
        for (int j = w; j <= C; j += QRW_THREADS / 64) {
            double f = 1.0;
            if (j < ce && a.cs) f = column_scale(a, t, j);
            else if (j >= ce && j < C && a.cs) f = a.cs[a.nr + so + (j - ce)];
            double *Y = X + (int64_t)j * M;
            for (int i = lane; i < M; i += 64) {
                double v = 0.0;
                if (j < C) {
                    if (i < me) {
                        const double x = (double)a.Rw[off + (int64_t)j * M + i];
                        bad += x - x;
                        v = (j >= ce || a.cs) ? (double)(T)(x * f) : x;
                    } else if (j >= ce && i == me + (j - ce)) {
                        v = (double)(T)(sqrt(a.lambda * a.scaling_e[so + (j - ce)]) * f);
                    }
                } else if (i < me) {
                    v = a.Rr[rq + i];
                    bad += v - v;
                }
                Y[i] = v;
            }
        }
        if (isnan(bad)) bad_sh = 1;   // every writer stores 1
        __syncthreads();
        bool failed = bad_sh != 0;
        const double tol = (double)M * (sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-07);
        for (int k = 0; k < de && !failed; k++) {
            double *P = X + (int64_t)(ce + k) * M;
            double pt = 0.0, ps = 0.0;
            for (int i = tid; i < M; i += QRW_THREADS) {
                const double x = P[i];
                if (i < k) pt += x * x;
                else ps += x * x;
            }
            pt = wave_sum(pt);
            ps = wave_sum(ps);
            if (lane == 0) { red[0][w] = pt; red[1][w] = ps; }
            __syncthreads();
            double s_top = 0.0, s_sub = 0.0;
            for (int q = 0; q < QRW_THREADS / 64; q++) { s_top += red[0][q]; s_sub += red[1][q]; }
            const double nrm = sqrt(s_sub), full = sqrt(s_top + s_sub);
            const double akk = P[k];
            const double beta = akk >= 0.0 ? -nrm : nrm;
            __syncthreads();   // every thread has read red and P[k]
            if (!(isfinite(beta) && fabs(beta) > tol * full)) { failed = true; break; }   // uniform
            const double scale = 1.0 / (akk - beta), tk = (beta - akk) / beta;
            for (int i = k + 1 + tid; i < M; i += QRW_THREADS) P[i] *= scale;   // v_k below row k (v_k[k] = 1)
            if (tid == 0) P[k] = beta;
            __syncthreads();
            // H_k on every other column: the later eliminated ones, the retained ones, the residual
            for (int j = w; j <= C; j += QRW_THREADS / 64) {
                if (j >= ce && j <= ce + k) continue;
                double *Y = X + (int64_t)j * M;
                const double yk = Y[k];
                double d = 0.0;
                for (int i = k + 1 + lane; i < M; i += 64) d += P[i] * Y[i];
                const double sj = (wave_sum(d) + yk) * tk;
                for (int i = k + 1 + lane; i < M; i += 64) Y[i] -= sj * P[i];
                if (lane == 0) Y[k] = yk - sj;
            }
            __syncthreads();
        }
        if (failed) {
            if (tid == 0) atomicMin(a.flag, (unsigned long long)(t + 1));
        } else {
            for (int j = w; j <= C; j += QRW_THREADS / 64) {
                const double *Y = X + (int64_t)j * M;
                const int k = j - ce;   // eliminated column k: U above and on the diagonal, zeros below
                for (int i = lane; i < M; i += 64) {
                    if (j < ce) a.F[off + (int64_t)j * M + i] = (T)Y[i];
                    else if (j < C) a.F[off + (int64_t)j * M + i] = i <= k ? (T)Y[i] : (T)0;
                    else a.Fr[rq + i] = Y[i];
                }
            }
        }
        __syncthreads();   // the workspace and bad_sh are reused by the next table
    }
}

// ---------------------------------------------------------------------------------------------------------
// Products with the reduced Jacobian J~ (the rows below U in every factored table; src/qr_reduced.jl:161-243,
// :484-512). Table t contributes n_t = M_t - top_t reduced rows at reduced_offsets[t]; its entry u holds retained
// block entry_retained[u] at table column entry_col[u].
// ---------------------------------------------------------------------------------------------------------
struct RedArgs {
    int64_t ntables, nrb;
    const int32_t *nrows, *top;
    const int64_t *raw_off, *res_off, *red_off;
    const int32_t *table_ptr, *entry_retained, *entry_col, *entry_table;
    const int32_t *retained_ptr, *retained_entry;
    const int32_t *r_dims;
    const int64_t *boff_r, *doff;
};

// y = J~ v (reduced_mul!, src/qr_reduced.jl:161-193): one wavefront per table, lane i its reduced row i; sums over
// the table's entries and their columns in the reference's order
template <class T>
__global__ __launch_bounds__(256) void k_red_mul(RedArgs a, const T *F, const double *v, double *y) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;
    const int M = a.nrows[t], top = a.top[t], n = M - top;
    for (int i = lane; i < n; i += 64) {
        double acc = 0.0;
        for (int u = a.table_ptr[t]; u < a.table_ptr[t + 1]; u++) {
            const int r = a.entry_retained[u];
            const int64_t cb = a.raw_off[t] + (int64_t)a.entry_col[u] * M + top + i;
            const double *vr = v + a.boff_r[r];
            for (int j = 0; j < a.r_dims[r]; j++) acc += (double)F[cb + (int64_t)j * M] * vr[j];
        }
        y[a.red_off[t] + i] = acc;
    }
}

// r~ (reduced_residual!, src/qr_reduced.jl:227-243): the residual rows below the top of each table
__global__ __launch_bounds__(256) void k_red_residual(RedArgs a, const double *Fr, double *rt) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;
    const int top = a.top[t], n = a.nrows[t] - top;
    for (int i = lane; i < n; i += 64) rt[a.red_off[t] + i] = Fr[a.res_off[t] + top + i];
}

// g = J~' w (reduced_tmul!, src/qr_reduced.jl:195-223) in two passes. k_red_tmul_part: one wavefront per table,
// lanes over the columns of its entries, writes the per-entry partials p[eoff[u] + j] = X_t(:, entry_col[u] + j)' w_t
// (eoff: prefix sums of the entry dims in entry order). k_red_gsum: one wavefront per retained block r adds the
// partials of the entries coupling r in retained_entry order: lane group q of the 64 / d_r groups takes the q-th
// contiguous run of that list, and the runs are then added in order, so the result is deterministic.
template <class T>
__global__ __launch_bounds__(256) void k_red_tmul_part(RedArgs a, const int64_t *__restrict__ eoff,
                                                       const T *__restrict__ F, const double *__restrict__ w,
                                                       double *__restrict__ gp) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;
    const int u0 = a.table_ptr[t], ne = a.table_ptr[t + 1] - u0;
    const int M = a.nrows[t], top = a.top[t], n = M - top;
    const T *Xt = F + a.raw_off[t] + top;
    const double *wt = w + a.red_off[t];
    const int64_t base = eoff[u0];
    const int ncol = (int)(eoff[u0 + ne] - base);
    for (int q = lane; q < ncol; q += 64) {
        int e = 0;
        while (e + 1 < ne && eoff[u0 + e + 1] - base <= q) e++;
        const T *x = Xt + (int64_t)(a.entry_col[u0 + e] + (q - (int)(eoff[u0 + e] - base))) * M;
        double acc = 0.0;
        int i = 0;
        for (; i + 8 <= n; i += 8) {
            double xv[8], wv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                xv[j] = (double)x[i + j];
                wv[j] = wt[i + j];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) acc += xv[j] * wv[j];
        }
        for (; i < n; i++) acc += (double)x[i] * wt[i];
        gp[base + q] = acc;
    }
}

__global__ __launch_bounds__(256) void k_red_gsum(RedArgs a, const int64_t *__restrict__ eoff,
                                                  const double *__restrict__ gp, double *__restrict__ g) {
    __shared__ double part[4][64];
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc = 0.0;
    int d = 1, ng = 0;
    if (r < a.nrb) {
        d = a.r_dims[r];
        ng = 64 / d;
        const int k0 = a.retained_ptr[r], L = a.retained_ptr[r + 1] - k0;
        const int grp = lane / d, j = lane - grp * d;
        if (grp < ng) {
            const int chunk = (L + ng - 1) / ng;
            const int kb = k0 + min(L, grp * chunk), ke = k0 + min(L, (grp + 1) * chunk);
            int k = kb;
            for (; k + 8 <= ke; k += 8) {
                double v[8];
#pragma unroll
                for (int q = 0; q < 8; q++) v[q] = gp[eoff[a.retained_entry[k + q]] + j];
#pragma unroll
                for (int q = 0; q < 8; q++) acc += v[q];
            }
            for (; k < ke; k++) acc += gp[eoff[a.retained_entry[k]] + j];
        }
    }
    part[wv][lane] = acc;
    __syncthreads();
    if (r < a.nrb && lane < d) {
        double sum = 0.0;
        for (int q = 0; q < ng; q++) sum += part[wv][q * d + lane];
        g[a.boff_r[r] + lane] = sum;
    }
}

// diagonal blocks of J~'J~ (reduced_diagonal_blocks!, src/qr_reduced.jl:484-512) in two passes. k_red_diag_part: one
// wavefront per table stages each entry's d columns of the reduced rows in LDS, 64 rows at a time (row stride 65, so
// the lanes of different columns hit different banks), and forms the upper pairs p = b (b + 1) / 2 + a (a <= b) of
// X_u' X_u into dp[dpo[u] + p] (dpo: prefix sums of d (d + 1) / 2 in entry order): for a table of at most 64 reduced
// rows on the FP64 matrix instruction (one 16 x 16 tile per entry), else lane p forms pair p by a dot product. k_red_dsum: one workgroup per
// retained block r adds the partials of the entries coupling r in retained_entry order -- thread (p, run) takes the
// run-th of 256 / npair contiguous runs of that list, and the runs are added in order -- and writes the full block,
// both triangles, column-major at doff[r].
__device__ inline void wave_sync() {   // LDS written by one lane of the wavefront, read by another
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double qd4 __attribute__((ext_vector_type(4)));

template <class T>
__global__ __launch_bounds__(256) void k_red_diag_part(RedArgs a, const int64_t *__restrict__ dpo,
                                                       const T *__restrict__ F, double *__restrict__ dp) {
    __shared__ double xs[4][16 * 65];
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;   // wavefront-level synchronisation only
    double *xw = xs[threadIdx.x >> 6];
    const int M = a.nrows[t], top = a.top[t], n = M - top;
    const T *Xt = F + a.raw_off[t] + top;
    const int ub = a.table_ptr[t], ue = a.table_ptr[t + 1];
    // pair p -> (a, b) of one entry's d x d block
    auto pairs = [&](int d, int *pa, int *pb) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int p = lane + 64 * k;
            int b = 0;
            while (b + 1 < d && (b + 1) * (b + 2) / 2 <= p) b++;
            pb[k] = b;
            pa[k] = min(p - b * (b + 1) / 2, b);
        }
    };
    if (n <= 64) {   // one row chunk: the next entry's columns load into registers while this one's dots run
        double nx[16];
        int d = 0, c0 = 0;
        if (ub < ue) {
            d = a.r_dims[a.entry_retained[ub]];
            c0 = a.entry_col[ub];
#pragma unroll
            for (int j = 0; j < 16; j++) nx[j] = (j < d && lane < n) ? (double)Xt[(int64_t)(c0 + j) * M + lane] : 0.0;
        }
        for (int u = ub; u < ue; u++) {
            const int dc = d;
#pragma unroll
            for (int j = 0; j < 16; j++)
                if (j < dc) xw[j * 65 + lane] = nx[j];
            wave_sync();
            if (u + 1 < ue) {
                d = a.r_dims[a.entry_retained[u + 1]];
                c0 = a.entry_col[u + 1];
#pragma unroll
                for (int j = 0; j < 16; j++) nx[j] = (j < d && lane < n) ? (double)Xt[(int64_t)(c0 + j) * M + lane] : 0.0;
            }
            // X_u'X_u on the FP64 matrix instruction, one 16 x 16 tile, k over the rows in steps of 4: lane
            // (li, lk) = (lane % 16, lane / 16) supplies A(li, k) = B(k, li) = X(k + lk, li) (columns past d and rows
            // past n read as zero), and accumulator v holds D(lk + 4 v, li)
            const int li = lane & 15, lk = lane >> 4;
            qd4 acc = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < n; k += 4) {
                const double xv = li < dc ? xw[li * 65 + k + lk] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, xv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int ra = lk + 4 * v;
                if (ra <= li && li < dc) dp[dpo[u] + li * (li + 1) / 2 + ra] = acc[v];
            }
            wave_sync();
        }
        return;
    }
    for (int u = ub; u < ue; u++) {
        const int d = a.r_dims[a.entry_retained[u]], c0 = a.entry_col[u];
        const int np = d * (d + 1) / 2;   // <= 136: three pairs per lane at most
        int pa[3], pb[3];
        pairs(d, pa, pb);
        double acc[3] = {0.0, 0.0, 0.0};
        for (int i0 = 0; i0 < n; i0 += 64) {
            const int i = i0 + lane;
            for (int j = 0; j < d; j++) xw[j * 65 + lane] = i < n ? (double)Xt[(int64_t)(c0 + j) * M + i] : 0.0;
            wave_sync();
            const int m = min(64, n - i0);
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (lane + 64 * k < np) {
                    const double *x = xw + pa[k] * 65, *y = xw + pb[k] * 65;
                    double s = 0.0;
#pragma unroll 8
                    for (int q = 0; q < m; q++) s += x[q] * y[q];
                    acc[k] += s;
                }
            wave_sync();
        }
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (lane + 64 * k < np) dp[dpo[u] + lane + 64 * k] = acc[k];
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_red_dsum(RedArgs a, const int64_t *__restrict__ dpo, const double *__restrict__ dp,
                                                  T *__restrict__ D) {
    __shared__ double part[256];
    const int r = blockIdx.x;
    const int d = a.r_dims[r], np = d * (d + 1) / 2;
    const int nrun = 256 / np;
    const int p = threadIdx.x % np, run = threadIdx.x / np;
    double acc = 0.0;
    if (run < nrun) {
        const int k0 = a.retained_ptr[r], L = a.retained_ptr[r + 1] - k0;
        const int len = (L + nrun - 1) / nrun;
        const int kb = k0 + min(L, run * len), ke = k0 + min(L, (run + 1) * len);
        int k = kb;
        for (; k + 8 <= ke; k += 8) {
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = dp[dpo[a.retained_entry[k + q]] + p];
#pragma unroll
            for (int q = 0; q < 8; q++) acc += v[q];
        }
        for (; k < ke; k++) acc += dp[dpo[a.retained_entry[k]] + p];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if ((int)threadIdx.x < np) {
        double sum = 0.0;
        for (int q = 0; q < nrun; q++) sum += part[q * np + threadIdx.x];
        int b = 0;
        while ((b + 1) * (b + 2) / 2 <= (int)threadIdx.x) b++;
        const int aa = threadIdx.x - b * (b + 1) / 2;
        T *Db = D + a.doff[r];
        Db[b * d + aa] = (T)sum;
        Db[aa * d + b] = (T)sum;
    }
}

// reduced_normal_dense! / reduced_normal_sparse! (src/qr_reduced.jl:373-480): the upper triangle of S = J~'J~ and
// g_S = J~'r~, by per-table Gram products. k_red_gram forms, for every table, the blocks X(:, cols(u1))' X(:, cols(u2))
// of each pair of its entries u1 <= u2 (X: the table's reduced rows) on the FP64 matrix instruction, each table read
// once, into a contribution buffer, and the per-entry partials X(:, cols(u))' r~. k_red_gram_sum adds the
// contributions of each block (r1, r2) of S in ascending table order -- the order in which the reference adds them to
// each entry -- and writes the whole block (the lower triangle of a diagonal block, and a block without
// contributions, as zeros); k_red_gsum adds the partials per retained block. Deterministic, no atomics.
#include "bls_gram.h"
#include "bls_fsum.h"

// Quantities of the full problem from the raw (or the factored) tables, all M rows of every column, accumulated in
// Float64 whatever the table type (src/qr_reduced.jl:574-645). Table t's raw columns are its entries' columns (entry
// u: r_dims[entry_retained[u]] columns from entry_col[u]) and, for an eliminated table, its d_e own columns last.
struct RawArgs {
    int64_t nr, neb;
    const int32_t *edims, *ncols;   // per table: d_e (0 for a free table), raw columns
    const int64_t *boff_e;          // per eliminated block, 0-based, into the eliminated part
};

// jtr! (MODE 0: column . residual, g = J'r) and jtj_diag! (MODE 1: column . column, diag(J'J)),
// src/qr_reduced.jl:574-596: one wavefront per table, lanes over its columns, each a run down the M rows in row
// order. The entries' columns write the per-entry partials gp[eoff[u] + j] (added per retained block by k_red_gsum,
// in retained_entry order, as sum_retained_columns! adds them); an eliminated table's own columns write
// out[nr + boffsets_e[t] + k] directly.
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_raw_colq(RedArgs a, RawArgs b, const int64_t *__restrict__ eoff,
                                                  const T *__restrict__ F, const double *__restrict__ Fr,
                                                  double *__restrict__ gp, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;
    const int u0 = a.table_ptr[t], ne = a.table_ptr[t + 1] - u0, M = a.nrows[t];
    const int de = b.edims[t];
    const T *Xt = F + a.raw_off[t];
    const double *rt = Fr + a.res_off[t];
    const int64_t base = eoff[u0];
    const int ncol = (int)(eoff[u0 + ne] - base);
    for (int q = lane; q < ncol + de; q += 64) {
        int col;
        if (q < ncol) {
            int e = 0;
            while (e + 1 < ne && eoff[u0 + e + 1] - base <= q) e++;
            col = a.entry_col[u0 + e] + (q - (int)(eoff[u0 + e] - base));
        } else {
            col = b.ncols[t] - de + (q - ncol);
        }
        const T *x = Xt + (int64_t)col * M;
        double acc = 0.0;
        int i = 0;
        for (; i + 8 <= M; i += 8) {   // eight loads in flight, added in row order
            double xv[8], wv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                xv[j] = (double)x[i + j];
                wv[j] = MODE == 1 ? xv[j] : rt[i + j];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) acc += xv[j] * wv[j];
        }
        for (; i < M; i++) {
            const double v = (double)x[i];
            acc += v * (MODE == 1 ? v : rt[i]);
        }
        if (q < ncol) gp[base + q] = acc;
        else out[b.nr + b.boff_e[t] + (q - ncol)] = acc;
    }
}

// ||J v||^2 (jtj_quadform, src/qr_reduced.jl:597-645): one wavefront per table, lane i its rows i, i + 64, ...; row
// i's y_i = sum over the entries in order, then the eliminated block, of J(i, cols) v(cols); the table's sum of y_i^2
// (the lanes' sums added in a fixed tree) into part[t], then a fixed-order sum over the tables
template <class T>
__global__ __launch_bounds__(256) void k_raw_quad(RedArgs a, RawArgs b, const T *__restrict__ F,
                                                  const double *__restrict__ v, double *__restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;
    const int u0 = a.table_ptr[t], u1 = a.table_ptr[t + 1], M = a.nrows[t];
    const int de = b.edims[t];
    const T *Xt = F + a.raw_off[t];
    double acc = 0.0;
    for (int i = lane; i < M; i += 64) {
        double yi = 0.0;
        for (int u = u0; u < u1; u++) {
            const int r = a.entry_retained[u];
            const T *x = Xt + (int64_t)a.entry_col[u] * M + i;
            const double *vr = v + a.boff_r[r];
            for (int j = 0; j < a.r_dims[r]; j++) yi += (double)x[(int64_t)j * M] * vr[j];
        }
        if (de) {
            const T *x = Xt + (int64_t)(b.ncols[t] - de) * M + i;
            const double *ve = v + b.nr + b.boff_e[t];
            for (int k = 0; k < de; k++) yi += (double)x[(int64_t)k * M] * ve[k];
        }
        acc += yi * yi;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) part[t] = acc;
}

// reduced_jacobian! (src/qr_reduced.jl:247-270, :334-370): J~ written out, sparse into the nzval of
// build_reduced_jacobian_pattern or dense (reduced_nrows x nr, column-major, zeroed first). One wavefront per table;
// lane q takes value q of the table's retained columns of reduced rows (column q / n, row q % n, n = M - top), so
// consecutive lanes read consecutive rows of a column and write consecutive entries of a run. Sparse: every column of
// retained block r holds h_r values, the runs of its tables in retained_entry order; entry u's run starts at jdst[u]
// in the first column of r and at jdst[u] + j h_r in column j. Plain copies: bitwise.
template <class T, bool SPARSE>
__global__ __launch_bounds__(256) void k_red_jacobian(RedArgs a, const int64_t *__restrict__ jdst,
                                                      const int64_t *__restrict__ jh, int64_t m,
                                                      const T *__restrict__ F, T *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= a.ntables) return;
    const int M = a.nrows[t], top = a.top[t], n = M - top;
    if (n <= 0) return;
    const T *Xt = F + a.raw_off[t] + top;
    for (int u = a.table_ptr[t]; u < a.table_ptr[t + 1]; u++) {
        const int r = a.entry_retained[u], d = a.r_dims[r];
        const T *x = Xt + (int64_t)a.entry_col[u] * M;
        for (int q = lane; q < d * n; q += 64) {
            const int j = q / n, i = q - j * n;
            const T v = x[(int64_t)j * M + i];
            if (SPARSE) out[jdst[u] + (int64_t)j * jh[r] + i] = v;
            else out[(int64_t)(a.boff_r[r] + j) * m + a.red_off[t] + i] = v;
        }
    }
}

template <class T> __global__ void k_fill_zero(int64_t n, T *p) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)0;
}

// backsubstitute! of the QR path (src/qr_reduced.jl:24-110): x_e = U^-1 (q - B x_r) per eliminated table, one
// lane per table: q the first d_e rows of the factored residual, B the first d_e rows of the retained columns, U the
// triangle of the eliminated columns; accumulated in Float64 whatever the table type.
constexpr int QB_LANES = 16;   // lanes per table in k_qr_backsub
template <class T>
__global__ __launch_bounds__(256) void k_qr_backsub(int64_t neb, const int32_t *nrows, const int32_t *ncols,
                                                    const int32_t *edims, const int64_t *raw_off,
                                                    const int64_t *res_off, const int64_t *boff_e,
                                                    const int32_t *table_ptr, const int32_t *entry_retained,
                                                    const int32_t *entry_col, const int32_t *r_dims,
                                                    const int64_t *boff_r, const T *F, const double *Fr,
                                                    const double *x_r, double *x_e) {
    // QB_LANES lanes per table: lane q adds B x_r over the entries q, q + QB_LANES, ... of the table (their columns'
    // first d_e rows); the lanes' partials are added in a fixed tree, and lane 0 solves U x_e = q - B x_r
    const int64_t e = (int64_t)blockIdx.x * (256 / QB_LANES) + threadIdx.x / QB_LANES;
    const int lane = threadIdx.x % QB_LANES;
    const bool valid = e < neb && edims[e] <= QR_MAXDE;   // larger blocks: k_qr_backsub_wide
    const int de = valid ? edims[e] : 0, M = valid ? nrows[e] : 0;
    const int64_t off = valid ? raw_off[e] : 0;
    double t[QR_MAXDE];
#pragma unroll
    for (int i = 0; i < QR_MAXDE; i++) t[i] = 0.0;
    const int u0 = valid ? table_ptr[e] : 0, u1 = valid ? table_ptr[e + 1] : 0;
    for (int u = u0 + lane; u < u1; u += QB_LANES) {
        const int r = entry_retained[u];
        const double *xr = x_r + boff_r[r];
        const T *cb = F + off + (int64_t)entry_col[u] * M;
        for (int j = 0; j < r_dims[r]; j++) {
            const double xj = xr[j];
#pragma unroll
            for (int i = 0; i < QR_MAXDE; i++)
                if (i < de) t[i] += (double)cb[(int64_t)j * M + i] * xj;
        }
    }
#pragma unroll
    for (int i = 0; i < QR_MAXDE; i++)
#pragma unroll
        for (int o = QB_LANES / 2; o >= 1; o >>= 1) t[i] += __shfl_xor(t[i], o, 64);
    if (!valid || lane != 0) return;
    const int ce = ncols[e] - de;
    for (int i = 0; i < de; i++) t[i] = Fr[res_off[e] + i] - t[i];
    double *xe = x_e + boff_e[e];
    for (int i = de - 1; i >= 0; i--) {
        double acc = t[i];
        for (int j = i + 1; j < de; j++) acc -= (double)F[off + (int64_t)(ce + j) * M + i] * xe[j];
        xe[i] = acc / (double)F[off + (int64_t)(ce + i) * M + i];
    }
}

// The same for the tables of d_e > QR_MAXDE: one workgroup per table (its index in `list`), t = q - B x_r in LDS
// (thread i its row i over the entries in order), then the triangle by thread 0.
template <class T>
__global__ __launch_bounds__(256) void k_qr_backsub_wide(const int32_t *list, int64_t nlist, const int32_t *nrows,
                                                         const int32_t *ncols, const int32_t *edims,
                                                         const int64_t *raw_off, const int64_t *res_off,
                                                         const int64_t *boff_e, const int32_t *table_ptr,
                                                         const int32_t *entry_retained, const int32_t *entry_col,
                                                         const int32_t *r_dims, const int64_t *boff_r, const T *F,
                                                         const double *Fr, const double *x_r, double *x_e) {
    extern __shared__ double tq[];
    for (int64_t li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int e = list[li];
        const int de = edims[e], M = nrows[e];
        const int64_t off = raw_off[e];
        for (int i = threadIdx.x; i < de; i += blockDim.x) {
            double acc = 0.0;
            for (int u = table_ptr[e]; u < table_ptr[e + 1]; u++) {
                const int r = entry_retained[u];
                const double *xr = x_r + boff_r[r];
                const T *cb = F + off + (int64_t)entry_col[u] * M;
                for (int j = 0; j < r_dims[r]; j++) acc += (double)cb[(int64_t)j * M + i] * xr[j];
            }
            tq[i] = Fr[res_off[e] + i] - acc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int ce = ncols[e] - de;
            double *xe = x_e + boff_e[e];
            for (int i = de - 1; i >= 0; i--) {
                double acc = tq[i];
                for (int j = i + 1; j < de; j++) acc -= (double)F[off + (int64_t)(ce + j) * M + i] * xe[j];
                xe[i] = acc / (double)F[off + (int64_t)(ce + i) * M + i];
            }
        }
        __syncthreads();
    }
}

__global__ void k_qr_fill_u64(unsigned long long *p, unsigned long long v) { *p = v; }
__global__ void k_qr_finalize(const unsigned long long *flag, int64_t *out) {
    const unsigned long long f = *flag;
    *out = f == (unsigned long long)NO_FAILURE ? 0 : (int64_t)f;
}

}  // namespace

struct bls_qr {
    int64_t nr = 0, nrb = 0, neb = 0, ntables = 0, raw_len = 0, res_len = 0;
    int max_de = 0;   // largest eliminated block dimension among the register kernel's tables
    std::vector<int32_t> bucket_tables[3];
    // eliminated tables of more than 64 rows or d_e > QR_MAXDE: k_qr_table_wide, with a Float64 workspace of
    // wide_grid x wide_wsz; those of d_e > QR_MAXDE also take k_qr_backsub_wide
    std::vector<int32_t> wide_tables, bigde_tables;
    DevArray<int32_t> d_wide, d_bigde;
    DevArray<double> d_wide_ws;
    int64_t wide_wsz = 0, wide_grid = 0;
    int bigde_max = 0;
    DevArray<int32_t> d_bucket[3];
    DevArray<int32_t> d_nrows, d_ncols, d_edims;
    DevArray<int64_t> d_raw_off, d_res_off, d_boff_e;
    DevArray<unsigned long long> d_flag;
    DevArray<int64_t> d_failed_out;
    Timing timing;
    // reduced-Jacobian structure (bls_qr_set_reduced)
    bool has_reduced = false;
    int64_t red_len = 0, dlen = 0, dp_len = 0;
    DevArray<int32_t> d_top, d_table_ptr, d_entry_retained, d_entry_col, d_entry_table, d_retained_ptr,
        d_retained_entry, d_r_dims;
    DevArray<int64_t> d_red_off, d_boff_r, d_doff;
    // the Gram formulation: per-table records in launch order (with offsets into the contribution buffer), and per
    // layout the S blocks with their contribution lists (ascending table order)
    bool rg_built[2] = {false, false};
    DevArray<int64_t> d_rg_ptr[2], d_rg_contrib[2];
    DevArray<RGBlock> d_rg_blocks[2];
    int64_t rg_nblk[2] = {0, 0}, rg_glen = 0;
    DevArray<double> d_rg_G;
    DevArray<RGTab> d_rg_tabs;                  // tables of the four launch classes, one after another
    DevArray<int32_t> d_rg_ecd;                 // per entry: (column - first column of its table) << 16 | dim
    DevArray<int64_t> d_eoff;                   // per-entry offsets of the partials of J~' w (prefix of entry dims)
    DevArray<double> d_gp;                      // the partials
    DevArray<int64_t> d_dpo;                    // per-entry offsets of the diagonal-block partials (d (d + 1) / 2 each)
    DevArray<double> d_dp;                      // the diagonal-block partials (allocated at the first call)
    DevArray<double> d_qpart;                   // ||J v||^2: per-table partials, then FSUM_CHUNKS chunk sums
    // build_reduced_jacobian_pattern's layout: per entry its run's start (0-based, first column of its block), per
    // retained block its column height; host copies for building and checking the pattern
    DevArray<int64_t> d_jdst, d_jh;
    int64_t jac_nnz = 0;
    std::vector<int64_t> h_jh, h_red_off;
    std::vector<int32_t> h_re, h_et;
    RGPlan rg;                                  // the Gram pass over those records
    std::vector<int32_t> h_tp, h_er, h_ec;   // table_ptr, entry_retained, entry_col (0-based)
    std::vector<int32_t> h_nrows, h_top;
    std::vector<int64_t> h_raw_off, h_res_off;   // 0-based
    std::vector<int64_t> h_bcp, h_brv, h_bro, h_ch;   // the sparse pattern (0-based)
    std::vector<int64_t> h_retained_ptr, h_S_colptr;   // host copies for the plans (0-based)
    std::vector<int32_t> h_r_dims;
    std::vector<int64_t> h_boff_r;
    // the sparse reduced-system layout (bls_qr_set_sparse)
    bool has_sparse = false;
    int64_t S_nnz = 0;
    DevArray<int64_t> d_S_colptr, d_block_row_offsets, d_col_heights;
    DevArray<int32_t> d_block_colptr, d_block_rowval;
    RedArgs red() const {
        RedArgs a{};
        a.ntables = ntables;
        a.nrb = nrb;
        a.nrows = d_nrows.ptr;
        a.top = d_top.ptr;
        a.raw_off = d_raw_off.ptr;
        a.res_off = d_res_off.ptr;
        a.red_off = d_red_off.ptr;
        a.table_ptr = d_table_ptr.ptr;
        a.entry_retained = d_entry_retained.ptr;
        a.entry_col = d_entry_col.ptr;
        a.entry_table = d_entry_table.ptr;
        a.retained_ptr = d_retained_ptr.ptr;
        a.retained_entry = d_retained_entry.ptr;
        a.r_dims = d_r_dims.ptr;
        a.boff_r = d_boff_r.ptr;
        a.doff = d_doff.ptr;
        return a;
    }
};

namespace {
std::vector<int64_t> sub1(const int64_t *p, int64_t n) {
    std::vector<int64_t> v(p, p + n);
    for (auto &x : v) x -= 1;
    return v;
}

template <class T>
int reduce_qr_async(bls_qr *h, T *F, double *Fr, const T *Rw, const double *Rr, const double *scaling_e,
                    double lambda, const double *cs, int64_t *d_failed, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (cs && !h->has_reduced)
        return fail(BLS_ERR_STATE, "reduce_qr with a conditioning scale needs the table entries (bls_qr_set_reduced)");
    if (!(std::isfinite(lambda) && lambda >= 0))
        return fail(BLS_ERR_ARG, "lambda must be finite and non-negative, got " + std::to_string(lambda));
    if ((h->raw_len && (!F || !Rw)) || (h->res_len && (!Fr || !Rr)) || (h->neb && !scaling_e) || !d_failed)
        return fail(BLS_ERR_ARG, "reduce_qr: a required array is NULL");
    h->timing.mark(0, s);
    k_qr_fill_u64<<<1, 1, 0, s>>>(h->d_flag.ptr, (unsigned long long)NO_FAILURE);
    QRArgs<T> a{};
    a.nrows = h->d_nrows.ptr;
    a.ncols = h->d_ncols.ptr;
    a.e_dims = h->d_edims.ptr;
    a.raw_off = h->d_raw_off.ptr;
    a.res_off = h->d_res_off.ptr;
    a.boff_e = h->d_boff_e.ptr;
    a.Rw = Rw;
    a.Rr = Rr;
    a.scaling_e = scaling_e;
    a.lambda = lambda;
    a.F = F;
    a.Fr = Fr;
    a.flag = h->d_flag.ptr;
    a.cs = cs;
    a.nr = h->nr;
    a.table_ptr = h->d_table_ptr.ptr;
    a.entry_retained = h->d_entry_retained.ptr;
    a.entry_col = h->d_entry_col.ptr;
    a.r_dims = h->d_r_dims.ptr;
    a.boff_r = h->d_boff_r.ptr;
    for (int b = 0; b < 3; b++) {
        a.tables = h->d_bucket[b].ptr;
        a.ntab = (int64_t)h->bucket_tables[b].size();
        if (!a.ntab) continue;
        // four wavefronts per workgroup, four tables per wavefront
        const unsigned grid = (unsigned)((a.ntab + 15) / 16);
        if (h->max_de <= 3) {
            if (b == 0) k_qr_tables<1, 3, T><<<grid, 256, 0, s>>>(a);
            else if (b == 1) k_qr_tables<2, 3, T><<<grid, 256, 0, s>>>(a);
            else k_qr_tables<4, 3, T><<<grid, 256, 0, s>>>(a);
        } else {
            if (b == 0) k_qr_tables<1, QR_MAXDE, T><<<grid, 256, 0, s>>>(a);
            else if (b == 1) k_qr_tables<2, QR_MAXDE, T><<<grid, 256, 0, s>>>(a);
            else k_qr_tables<4, QR_MAXDE, T><<<grid, 256, 0, s>>>(a);
        }
    }
    if (!h->wide_tables.empty()) {
        a.tables = h->d_wide.ptr;
        a.ntab = (int64_t)h->wide_tables.size();
        k_qr_table_wide<T><<<(unsigned)h->wide_grid, QRW_THREADS, 0, s>>>(a, h->d_wide_ws.ptr, h->wide_wsz);
    }
    h->timing.mark(1, s);
    k_qr_finalize<<<1, 1, 0, s>>>(h->d_flag.ptr, d_failed);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
}  // namespace

extern "C" int bls_qr_create(int64_t nr, int64_t nrblocks, int64_t neblocks, const int64_t *e_dims,
                             const int64_t *boffsets_e, int64_t ntables, const int64_t *table_nrows,
                             const int64_t *table_ncols, const int64_t *raw_offsets, const int64_t *residual_offsets,
                             bls_qr **out) {
    if (!out || nr < 0 || nrblocks < 0 || neblocks < 0 || ntables < neblocks) return fail(BLS_ERR_ARG, "bls_qr_create: bad arguments");
    *out = nullptr;
    if (ntables > INT32_MAX) return fail(BLS_ERR_UNSUPPORTED, "more than 2^31 tables");
    auto *h = new bls_qr();
    h->nr = nr;
    h->nrb = nrblocks;
    h->neb = neblocks;
    h->ntables = ntables;
    std::vector<int32_t> nrows(ntables), ncols(ntables), edims(ntables, 0);
    std::vector<int64_t> boff(ntables, 0);
    for (int64_t t = 0; t < ntables; t++) {
        nrows[t] = (int32_t)table_nrows[t];
        ncols[t] = (int32_t)(table_ncols[t] - 1);
        if (t < neblocks) {
            edims[t] = (int32_t)e_dims[t];
            boff[t] = boffsets_e[t] - 1;
            if (e_dims[t] < 1 || e_dims[t] > QRW_MAXDE) {
                delete h;
                return fail(BLS_ERR_UNSUPPORTED, "eliminated block dimension outside 1.." + std::to_string(QRW_MAXDE) +
                                                     " on the device QR path");
            }
            if (e_dims[t] > QR_MAXDE) {
                h->bigde_tables.push_back((int32_t)t);
                h->bigde_max = std::max<int>(h->bigde_max, (int)e_dims[t]);
            }
            if (table_nrows[t] > 64 || e_dims[t] > QR_MAXDE) {
                h->wide_tables.push_back((int32_t)t);
                h->wide_wsz = std::max<int64_t>(h->wide_wsz, (int64_t)table_nrows[t] * table_ncols[t]);
                continue;
            }
            h->max_de = std::max<int>(h->max_de, (int)e_dims[t]);
        }
        const int b = t >= neblocks ? 0 : (table_nrows[t] <= 16 ? 0 : table_nrows[t] <= 32 ? 1 : 2);
        h->bucket_tables[b].push_back((int32_t)t);
    }
    // within a bucket, tables with the most columns first: the four tables of a wavefront run in lockstep, so
    // neighbours of similar width waste little, and the dispatcher fills the tail with the narrow ones
    for (int b = 0; b < 3; b++)
        std::stable_sort(h->bucket_tables[b].begin(), h->bucket_tables[b].end(),
                         [&](int32_t x, int32_t y) { return ncols[x] > ncols[y]; });
    h->raw_len = raw_offsets[ntables] - 1;
    h->res_len = residual_offsets[ntables] - 1;
    int rc;
    h->h_raw_off = sub1(raw_offsets, ntables + 1);
    h->h_res_off = sub1(residual_offsets, ntables + 1);
    if ((rc = h->d_nrows.upload(nrows)) || (rc = h->d_ncols.upload(ncols)) || (rc = h->d_edims.upload(edims)) ||
        (rc = h->d_raw_off.upload(h->h_raw_off)) || (rc = h->d_res_off.upload(h->h_res_off)) ||
        (rc = h->d_boff_e.upload(boff)) ||
        (rc = h->d_flag.alloc(1)) || (rc = h->d_failed_out.alloc(1))) {
        delete h;
        return rc;
    }
    for (int b = 0; b < 3; b++)
        if ((rc = h->d_bucket[b].upload(h->bucket_tables[b]))) { delete h; return rc; }
    if (!h->wide_tables.empty()) {   // M x (C + 1) doubles per workgroup (table_ncols counts the residual)
        h->wide_grid = std::min<int64_t>((int64_t)h->wide_tables.size(),
                                         std::max<int64_t>(1, std::min<int64_t>(1024, (1ll << 27) / h->wide_wsz)));
        if ((rc = h->d_wide.upload(h->wide_tables)) || (rc = h->d_bigde.upload(h->bigde_tables)) ||
            (rc = h->d_wide_ws.alloc((size_t)(h->wide_grid * h->wide_wsz)))) {
            delete h;
            return rc;
        }
    }
    h->h_nrows = nrows;
    *out = h;
    return BLS_OK;
}

extern "C" void bls_qr_destroy(bls_qr *h) { delete h; }

extern "C" int bls_qr_set_reduced(bls_qr *h, const int64_t *table_top, const int64_t *reduced_offsets,
                                  const int64_t *table_ptr, int64_t nentries, const int64_t *entry_retained,
                                  const int64_t *entry_col, const int64_t *entry_table, const int64_t *retained_ptr,
                                  const int64_t *retained_entry, const int64_t *r_dims, const int64_t *boffsets_r,
                                  const int64_t *doffsets) {
    if (!h || !table_top || !reduced_offsets || !table_ptr || nentries < 0 || !r_dims || !boffsets_r || !doffsets ||
        (nentries && (!entry_retained || !entry_col || !entry_table || !retained_ptr || !retained_entry)))
        return fail(BLS_ERR_ARG, "bls_qr_set_reduced: bad arguments");
    if (nentries > INT32_MAX) return fail(BLS_ERR_UNSUPPORTED, "more than 2^31 table entries");
    const int64_t nt = h->ntables, nrb = h->nrb;
    std::vector<int32_t> top(nt), tp(nt + 1), er(nentries), ec(nentries), et(nentries), rp(nrb + 1), re(nentries),
        rd(nrb);
    std::vector<int64_t> ro(nt + 1), bo(nrb), dof(nrb);
    for (int64_t t = 0; t < nt; t++) top[t] = (int32_t)table_top[t];
    for (int64_t t = 0; t <= nt; t++) { tp[t] = (int32_t)(table_ptr[t] - 1); ro[t] = reduced_offsets[t] - 1; }
    for (int64_t u = 0; u < nentries; u++) {
        er[u] = (int32_t)(entry_retained[u] - 1);
        ec[u] = (int32_t)(entry_col[u] - 1);
        et[u] = (int32_t)(entry_table[u] - 1);
        re[u] = (int32_t)(retained_entry[u] - 1);
        if (er[u] < 0 || er[u] >= nrb || et[u] < 0 || et[u] >= nt || re[u] < 0 || re[u] >= nentries)
            return fail(BLS_ERR_ARG, "bls_qr_set_reduced: entry index out of range");
    }
    for (int64_t r = 0; r <= nrb; r++) rp[r] = (int32_t)(retained_ptr[r] - 1);
    int64_t dl = 0;
    for (int64_t r = 0; r < nrb; r++) {
        rd[r] = (int32_t)r_dims[r];
        bo[r] = boffsets_r[r] - 1;
        dof[r] = doffsets[r] - 1;
        if (rd[r] < 1 || rd[r] > 16) return fail(BLS_ERR_UNSUPPORTED, "retained block dimension outside 1..16");
        dl = std::max<int64_t>(dl, dof[r] + (int64_t)rd[r] * rd[r]);
    }
    int rc;
    if ((rc = h->d_top.upload(top)) || (rc = h->d_table_ptr.upload(tp)) || (rc = h->d_entry_retained.upload(er)) ||
        (rc = h->d_entry_col.upload(ec)) || (rc = h->d_entry_table.upload(et)) || (rc = h->d_retained_ptr.upload(rp)) ||
        (rc = h->d_retained_entry.upload(re)) || (rc = h->d_r_dims.upload(rd)) || (rc = h->d_red_off.upload(ro)) ||
        (rc = h->d_boff_r.upload(bo)) || (rc = h->d_doff.upload(dof)))
        return rc;
    std::vector<int64_t> eoff(nentries + 1, 0);
    for (int64_t u = 0; u < nentries; u++) eoff[u + 1] = eoff[u] + rd[er[u]];
    if ((rc = h->d_eoff.upload(eoff))) return rc;
    if ((int64_t)h->d_gp.n < eoff[nentries] && (rc = h->d_gp.alloc((size_t)std::max<int64_t>(1, eoff[nentries])))) return rc;
    std::vector<int64_t> dpo(nentries + 1, 0);
    for (int64_t u = 0; u < nentries; u++) dpo[u + 1] = dpo[u] + (int64_t)rd[er[u]] * (rd[er[u]] + 1) / 2;
    if ((rc = h->d_dpo.upload(dpo))) return rc;
    if ((rc = h->d_qpart.alloc((size_t)(nt + FSUM_CHUNKS)))) return rc;
    {   // build_reduced_jacobian_pattern (src/qr_reduced.jl:289-310): columns in retained-block order, each column of
        // block r holding the reduced rows of r's tables in retained_entry order
        std::vector<int64_t> jh(nrb), jdst(nentries);
        int64_t cp = 0;
        for (int64_t r = 0; r < nrb; r++) {
            int64_t pos = 0;
            for (int k = rp[r]; k < rp[r + 1]; k++) {
                const int32_t u = re[k], t = et[u];
                jdst[u] = cp + pos;
                pos += h->h_nrows[t] - top[t];
            }
            jh[r] = pos;
            cp += pos * rd[r];
        }
        if ((rc = h->d_jdst.upload(jdst)) || (rc = h->d_jh.upload(jh))) return rc;
        h->jac_nnz = cp;
        h->h_jh = jh;
        h->h_red_off = ro;
        h->h_re = re;
        h->h_et = et;
    }
    h->dp_len = dpo[nentries];
    h->red_len = ro[nt];
    h->dlen = dl;
    h->h_retained_ptr.assign(rp.begin(), rp.end());
    h->h_tp = tp;
    h->h_top = top;
    h->h_er = er;
    h->h_ec = ec;
    h->h_r_dims = rd;
    h->h_boff_r = bo;
    h->has_reduced = true;
    return BLS_OK;
}

static int reduced_ready(bls_qr *h) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (!h->has_reduced) return fail(BLS_ERR_STATE, "no reduced-Jacobian structure (bls_qr_set_reduced)");
    return BLS_OK;
}

namespace {
template <class T> int reduced_mul(bls_qr *h, double *y, const T *F, const double *v, hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!y || !F || !v) return fail(BLS_ERR_ARG, "reduced_mul: a required array is NULL");
    if (h->ntables) k_red_mul<T><<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), F, v, y);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

RawArgs raw_args(const bls_qr *h) {
    RawArgs b{};
    b.nr = h->nr;
    b.neb = h->neb;
    b.edims = h->d_edims.ptr;
    b.ncols = h->d_ncols.ptr;
    b.boff_e = h->d_boff_e.ptr;
    return b;
}

// jtr! (MODE 0) and jtj_diag! (MODE 1) into out = [out_r; out_e] (length nr + sum(e_dims))
template <class T, int MODE> int raw_colq(bls_qr *h, double *out, const T *F, const double *r, hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!out || (h->raw_len && !F) || (MODE == 0 && h->res_len && !r))
        return fail(BLS_ERR_ARG, MODE ? "jtj_diag: a required array is NULL" : "jtr: a required array is NULL");
    if (h->ntables)
        k_raw_colq<T, MODE><<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), raw_args(h), h->d_eoff.ptr, F, r,
                                                                            h->d_gp.ptr, out);
    if (h->nrb) k_red_gsum<<<(unsigned)((h->nrb + 3) / 4), 256, 0, s>>>(h->red(), h->d_eoff.ptr, h->d_gp.ptr, out);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class T> int raw_quad(bls_qr *h, const T *F, const double *v, double *result, hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!result || !v || (h->raw_len && !F)) return fail(BLS_ERR_ARG, "jtj_quadform: a required array is NULL");
    if (h->ntables) k_raw_quad<T><<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), raw_args(h), F, v, h->d_qpart.ptr);
    fsum_launch(h->ntables, h->d_qpart.ptr, h->d_qpart.ptr + h->ntables, result, s);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class T> int reduced_tmul(bls_qr *h, double *g, const T *F, const double *w, hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!g || !F || !w) return fail(BLS_ERR_ARG, "reduced_tmul: a required array is NULL");
    if (h->ntables)
        k_red_tmul_part<T><<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), h->d_eoff.ptr, F, w, h->d_gp.ptr);
    if (h->nrb) k_red_gsum<<<(unsigned)((h->nrb + 3) / 4), 256, 0, s>>>(h->red(), h->d_eoff.ptr, h->d_gp.ptr, g);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class T> int reduced_diag(bls_qr *h, T *D, const T *F, hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!D || !F) return fail(BLS_ERR_ARG, "reduced_diagonal_blocks: a required array is NULL");
    if ((int64_t)h->d_dp.n < h->dp_len && (rc = h->d_dp.alloc((size_t)std::max<int64_t>(1, h->dp_len)))) return rc;
    if (h->ntables)
        k_red_diag_part<T><<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), h->d_dpo.ptr, F, h->d_dp.ptr);
    if (h->nrb) k_red_dsum<T><<<(unsigned)h->nrb, 256, 0, s>>>(h->red(), h->d_dpo.ptr, h->d_dp.ptr, D);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class T>
int reduce_qr_sync(bls_qr *h, T *factored, double *fres, const T *raw, const double *r, const double *scaling_e,
                   double lambda, const double *s, int64_t *failed_block, hipStream_t st) {
    int rc = reduce_qr_async<T>(h, factored, fres, raw, r, scaling_e, lambda, s, h ? h->d_failed_out.ptr : nullptr, st);
    if (rc) return rc;
    int64_t fb = 0;
    BLS_HIP(hipMemcpyAsync(&fb, h->d_failed_out.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    BLS_HIP(hipStreamSynchronize(st));
    if (h->timing.enabled) (void)hipEventElapsedTime(&h->timing.ms[0], h->timing.ev[0], h->timing.ev[1]);
    if (failed_block) *failed_block = fb;
    return fb ? BLS_FAILED : BLS_OK;
}
}  // namespace

namespace {
template <class T> int qr_backsub(bls_qr *h, double *x_e, const double *x_r, const T *F, const double *Fr,
                                  hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (h->neb && (!x_e || !x_r || !F || !Fr)) return fail(BLS_ERR_ARG, "backsubstitute: a required array is NULL");
    if (h->neb)
        k_qr_backsub<T><<<grid_for(h->neb, 256 / QB_LANES), 256, 0, s>>>(h->neb, h->d_nrows.ptr, h->d_ncols.ptr, h->d_edims.ptr,
                                                             h->d_raw_off.ptr, h->d_res_off.ptr, h->d_boff_e.ptr,
                                                             h->d_table_ptr.ptr, h->d_entry_retained.ptr,
                                                             h->d_entry_col.ptr, h->d_r_dims.ptr, h->d_boff_r.ptr, F,
                                                             Fr, x_r, x_e);
    if (!h->bigde_tables.empty())
        k_qr_backsub_wide<T><<<(unsigned)std::min<size_t>(h->bigde_tables.size(), 4096), 256,
                               (size_t)h->bigde_max * sizeof(double), s>>>(
            h->d_bigde.ptr, (int64_t)h->bigde_tables.size(), h->d_nrows.ptr, h->d_ncols.ptr, h->d_edims.ptr,
            h->d_raw_off.ptr, h->d_res_off.ptr, h->d_boff_e.ptr, h->d_table_ptr.ptr, h->d_entry_retained.ptr,
            h->d_entry_col.ptr, h->d_r_dims.ptr, h->d_boff_r.ptr, F, Fr, x_r, x_e);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
}  // namespace

extern "C" int bls_qr_set_sparse(bls_qr *h, int64_t nnz, const int64_t *S_colptr, int64_t nnz_blocks,
                                 const int64_t *block_colptr, const int64_t *block_rowval,
                                 const int64_t *block_row_offsets, const int64_t *col_heights,
                                 const int64_t *table_ptr, const int64_t *entry_retained) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (nnz < 0 || nnz_blocks < 0 || nnz_blocks > INT32_MAX || !S_colptr || !block_colptr || !col_heights ||
        (nnz_blocks && (!block_rowval || !block_row_offsets)) || !table_ptr)
        return fail(BLS_ERR_ARG, "bls_qr_set_sparse: bad arguments");
    const int64_t nrb = h->nrb, nt = h->ntables;
    std::vector<int32_t> bcp(nrb + 1), brv(nnz_blocks);
    for (int64_t r = 0; r <= nrb; r++) bcp[r] = (int32_t)(block_colptr[r] - 1);
    for (int64_t q = 0; q < nnz_blocks; q++) brv[q] = (int32_t)(block_rowval[q] - 1);
    // every pair of entries u1 <= u2 of a table needs the block (r1, r2) in the pattern (the reference throws
    // throw_missing_block otherwise)
    for (int64_t t = 0; t < nt; t++)
        for (int64_t u2 = table_ptr[t] - 1; u2 < table_ptr[t + 1] - 1; u2++) {
            const int32_t r2 = (int32_t)(entry_retained[u2] - 1);
            for (int64_t u1 = table_ptr[t] - 1; u1 <= u2; u1++) {
                const int32_t r1 = (int32_t)(entry_retained[u1] - 1);
                if (r2 < 0 || r2 >= nrb || !std::binary_search(brv.begin() + bcp[r2], brv.begin() + bcp[r2 + 1], r1))
                    return fail(BLS_ERR_ARG, "The sparse pattern has no block (" + std::to_string(r1 + 1) + ", " +
                                                 std::to_string(r2 + 1) + ")");
            }
        }
    std::vector<int64_t> Scp(S_colptr, S_colptr + h->nr + 1), bro(block_row_offsets, block_row_offsets + nnz_blocks),
        ch(col_heights, col_heights + nrb);
    for (auto &x : Scp) x -= 1;
    if ((rc = h->d_S_colptr.upload(Scp)) || (rc = h->d_block_colptr.upload(bcp)) || (rc = h->d_block_rowval.upload(brv)) ||
        (rc = h->d_block_row_offsets.upload(bro)) || (rc = h->d_col_heights.upload(ch)))
        return rc;
    h->S_nnz = nnz;
    h->h_S_colptr = Scp;
    h->h_bcp.assign(bcp.begin(), bcp.end());
    h->h_brv.assign(brv.begin(), brv.end());
    h->h_bro = bro;
    h->h_ch = ch;
    h->rg_built[0] = false;
    h->has_sparse = true;
    return BLS_OK;
}

namespace {
// The plan of the Gram formulation for one layout: per-table offsets of the pair blocks, the blocks of S with their
// write addresses, and each block's contributions in ascending table order.
int build_gram_plan(bls_qr *h, int dense) {
    const int m = dense ? 1 : 0;
    const int64_t nt = h->ntables, nrb = h->nrb;
    const std::vector<int32_t> &tp = h->h_tp, &er = h->h_er, &ec = h->h_ec, &rd = h->h_r_dims;
    const std::vector<int64_t> &bo = h->h_boff_r;
    std::vector<int64_t> goff(nt + 1);
    int64_t acc = 0;
    for (int64_t t = 0; t < nt; t++) {
        goff[t] = acc;
        const int ne = tp[t + 1] - tp[t];
        if (ne > RG_MAXE)
            return fail(BLS_ERR_UNSUPPORTED, "table " + std::to_string(t + 1) + " holds " + std::to_string(ne) +
                                                 " retained blocks; reduced_normal on the device takes up to 64");
        if (ne && ec[tp[t] + ne - 1] + rd[er[tp[t] + ne - 1]] - ec[tp[t]] > RG_MAXW)
            return fail(BLS_ERR_UNSUPPORTED, "table " + std::to_string(t + 1) +
                                                 ": its entries span more than 1024 columns; reduced_normal on the device "
                                                 "takes up to 1024");
        for (int i = 1; i < ne; i++)   // pair blocks land in the upper triangle; the last entry ends the columns
            if (er[tp[t] + i] <= er[tp[t] + i - 1] || ec[tp[t] + i] <= ec[tp[t] + i - 1])
                return fail(BLS_ERR_ARG, "table " + std::to_string(t + 1) +
                                             ": entries not in strictly ascending retained-block and column order");
        for (int i2 = 0; i2 < ne; i2++)
            for (int i1 = 0; i1 <= i2; i1++) acc += (int64_t)rd[er[tp[t] + i1]] * rd[er[tp[t] + i2]];
    }
    goff[nt] = acc;
    // one record per table with entries, ordered into the launch classes
    std::vector<int64_t> eoff(er.size() + 1, 0);
    for (size_t u = 0; u < er.size(); u++) eoff[u + 1] = eoff[u] + rd[er[u]];
    std::vector<RGTab> recs, tabs;
    for (int64_t t = 0; t < nt; t++) {
        const int u0 = tp[t], ne = tp[t + 1] - u0;
        if (!ne) continue;
        RGTab g;
        g.raw_off = h->h_raw_off[t];
        g.res_off = h->h_res_off[t];
        g.goff = goff[t];
        g.eoff = eoff[u0];
        g.M = h->h_nrows[t];
        g.top = h->h_top[t];
        g.n = g.M - g.top;
        g.u0 = u0;
        g.ne = ne;
        g.c0 = ec[u0];
        g.W = ec[u0 + ne - 1] + rd[er[u0 + ne - 1]] - ec[u0];
        g.pad = 0;
        recs.push_back(g);
    }
    int rc0;
    if ((rc0 = rg_classify(recs, tabs, h->rg))) return rc0;
    std::vector<int32_t> ecd(er.size());
    for (int64_t t = 0; t < nt; t++)
        for (int u = tp[t]; u < tp[t + 1]; u++) ecd[u] = ((ec[u] - ec[tp[t]]) << 16) | rd[er[u]];
    if ((rc0 = h->d_rg_tabs.upload(tabs)) || (rc0 = h->d_rg_ecd.upload(ecd))) return rc0;
    h->rg.tabs = h->d_rg_tabs.ptr;
    h->rg.ecd = h->d_rg_ecd.ptr;
    std::vector<RGBlock> blocks;
    if (dense) {
        blocks.reserve((size_t)(nrb * (nrb + 1) / 2));
        for (int64_t r2 = 0; r2 < nrb; r2++)
            for (int64_t r1 = 0; r1 <= r2; r1++)
                blocks.push_back(RGBlock{bo[r1] + bo[r2] * h->nr, h->nr, rd[r1], rd[r2], r1 == r2, 0});
    } else {
        for (int64_t r2 = 0; r2 < nrb; r2++)
            for (int64_t q = h->h_bcp[r2]; q < h->h_bcp[r2 + 1]; q++) {
                const int64_t r1 = h->h_brv[q];
                blocks.push_back(RGBlock{h->h_S_colptr[bo[r2]] + h->h_bro[q], h->h_ch[r2], rd[r1], rd[r2], r1 == r2, 0});
            }
    }
    const int64_t nblk = (int64_t)blocks.size();
    auto block_of = [&](int64_t r1, int64_t r2) -> int64_t {
        if (dense) return r2 * (r2 + 1) / 2 + r1;
        const int64_t *b0 = h->h_brv.data() + h->h_bcp[r2], *b1 = h->h_brv.data() + h->h_bcp[r2 + 1];
        return (int64_t)(std::lower_bound(b0, b1, r1) - h->h_brv.data());   // present: checked by bls_qr_set_sparse
    };
    std::vector<int64_t> ptr(nblk + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
        std::vector<int64_t> pos;
        std::vector<int64_t> contrib;
        if (pass == 1) {
            for (int64_t b = 0; b < nblk; b++) ptr[b + 1] += ptr[b];
            pos.assign(ptr.begin(), ptr.end() - 1);
            contrib.resize((size_t)ptr[nblk]);
        }
        for (int64_t t = 0; t < nt; t++) {
            const int ne = tp[t + 1] - tp[t];
            int64_t off = goff[t];
            for (int i2 = 0; i2 < ne; i2++)
                for (int i1 = 0; i1 <= i2; i1++) {
                    const int64_t r1 = er[tp[t] + i1], r2 = er[tp[t] + i2];
                    const int64_t b = block_of(r1, r2);
                    if (pass == 0) ptr[b + 1]++;
                    else contrib[(size_t)pos[b]++] = off;
                    off += (int64_t)rd[r1] * rd[r2];
                }
        }
        if (pass == 1) {
            int rc;
            if ((rc = h->d_rg_ptr[m].upload(ptr)) || (rc = h->d_rg_contrib[m].upload(contrib))) return rc;
        }
    }
    int rc;
    if ((rc = h->d_rg_blocks[m].upload(blocks))) return rc;
    if (acc > h->rg_glen) {
        if ((rc = h->d_rg_G.alloc((size_t)acc))) return rc;
        h->rg_glen = acc;
    }
    h->rg_nblk[m] = nblk;
    h->rg_built[m] = true;
    return BLS_OK;
}

template <class T> int reduced_normal(bls_qr *h, int dense, T *S, double *g, const T *F, const double *Fr,
                                      hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!dense && !h->has_sparse) return fail(BLS_ERR_STATE, "no sparse layout (bls_qr_set_sparse)");
    if (h->nrb && (!S || !g || !F || !Fr)) return fail(BLS_ERR_ARG, "reduced_normal: a required array is NULL");
    const int m = dense ? 1 : 0;
    if (!h->rg_built[m] && (rc = build_gram_plan(h, dense))) return rc;
    if (dense && h->nr) k_rg_fill<T><<<std::min<int64_t>(grid_for(h->nr * h->nr, 256), 4096), 256, 0, s>>>(S, h->nr * h->nr);
    if ((rc = rg_launch<T>(h->rg, F, Fr, h->d_rg_G.ptr, h->d_gp.ptr, s))) return rc;
    if (h->rg_nblk[m])
        k_red_gram_sum<T><<<(unsigned)h->rg_nblk[m], 256, 0, s>>>(h->rg_nblk[m], h->d_rg_blocks[m].ptr,
                                                                            h->d_rg_ptr[m].ptr, h->d_rg_contrib[m].ptr,
                                                                            h->d_rg_G.ptr, S);
    // g_S = J~' r~ from the per-entry partials the Gram kernel wrote (fixed-order sums)
    if (h->nrb) k_red_gsum<<<(unsigned)((h->nrb + 3) / 4), 256, 0, s>>>(h->red(), h->d_eoff.ptr, h->d_gp.ptr, g);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

}  // namespace

extern "C" int bls_reduced_normal_sparse_f64(bls_qr *h, double *S_nzval, double *g_S, const double *factored,
                                             const double *fres, void *stream) {
    return reduced_normal<double>(h, 0, S_nzval, g_S, factored, fres, (hipStream_t)stream);
}
extern "C" int bls_reduced_normal_sparse_f32(bls_qr *h, float *S_nzval, double *g_S, const float *factored,
                                             const double *fres, void *stream) {
    return reduced_normal<float>(h, 0, S_nzval, g_S, factored, fres, (hipStream_t)stream);
}
extern "C" int bls_reduced_normal_dense_f64(bls_qr *h, double *S, double *g_S, const double *factored,
                                            const double *fres, void *stream) {
    return reduced_normal<double>(h, 1, S, g_S, factored, fres, (hipStream_t)stream);
}
extern "C" int bls_reduced_normal_dense_f32(bls_qr *h, float *S, double *g_S, const float *factored,
                                            const double *fres, void *stream) {
    return reduced_normal<float>(h, 1, S, g_S, factored, fres, (hipStream_t)stream);
}

extern "C" int bls_qr_backsubstitute_f64(bls_qr *h, double *x_e, const double *x_r, const double *factored,
                                         const double *fres, void *stream) {
    return qr_backsub<double>(h, x_e, x_r, factored, fres, (hipStream_t)stream);
}
extern "C" int bls_qr_backsubstitute_f32(bls_qr *h, double *x_e, const double *x_r, const float *factored,
                                         const double *fres, void *stream) {
    return qr_backsub<float>(h, x_e, x_r, factored, fres, (hipStream_t)stream);
}

extern "C" int bls_reduced_mul_f64(bls_qr *h, double *y, const double *F, const double *v, void *stream) {
    return reduced_mul<double>(h, y, F, v, (hipStream_t)stream);
}
extern "C" int bls_reduced_mul_f32(bls_qr *h, double *y, const float *F, const double *v, void *stream) {
    return reduced_mul<float>(h, y, F, v, (hipStream_t)stream);
}
extern "C" int bls_reduced_tmul_f64(bls_qr *h, double *g, const double *F, const double *w, void *stream) {
    return reduced_tmul<double>(h, g, F, w, (hipStream_t)stream);
}
extern "C" int bls_reduced_tmul_f32(bls_qr *h, double *g, const float *F, const double *w, void *stream) {
    return reduced_tmul<float>(h, g, F, w, (hipStream_t)stream);
}

extern "C" int bls_jtr_f64(bls_qr *h, double *g, const double *F, const double *r, void *stream) {
    return raw_colq<double, 0>(h, g, F, r, (hipStream_t)stream);
}
extern "C" int bls_jtr_f32(bls_qr *h, double *g, const float *F, const double *r, void *stream) {
    return raw_colq<float, 0>(h, g, F, r, (hipStream_t)stream);
}
extern "C" int bls_jtj_diag_f64(bls_qr *h, double *d, const double *F, void *stream) {
    return raw_colq<double, 1>(h, d, F, nullptr, (hipStream_t)stream);
}
extern "C" int bls_jtj_diag_f32(bls_qr *h, double *d, const float *F, void *stream) {
    return raw_colq<float, 1>(h, d, F, nullptr, (hipStream_t)stream);
}
extern "C" int bls_jtj_quadform_f64(bls_qr *h, const double *F, const double *v, double *result, void *stream) {
    return raw_quad<double>(h, F, v, result, (hipStream_t)stream);
}
extern "C" int bls_jtj_quadform_f32(bls_qr *h, const float *F, const double *v, double *result, void *stream) {
    return raw_quad<float>(h, F, v, result, (hipStream_t)stream);
}

extern "C" int bls_reduced_jacobian_nnz(bls_qr *h, int64_t *nnz) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!nnz) return fail(BLS_ERR_ARG, "bls_reduced_jacobian_nnz: nnz is NULL");
    *nnz = h->jac_nnz;
    return BLS_OK;
}

extern "C" int bls_build_reduced_jacobian_pattern(bls_qr *h, int64_t *colptr, int64_t *rowval) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!colptr || (h->jac_nnz && !rowval)) return fail(BLS_ERR_ARG, "bls_build_reduced_jacobian_pattern: bad arguments");
    int64_t p = 0, c = 0;
    colptr[0] = 1;
    for (int64_t r = 0; r < h->nrb; r++)
        for (int j = 0; j < h->h_r_dims[r]; j++, c++) {
            for (int64_t k = h->h_retained_ptr[r]; k < h->h_retained_ptr[r + 1]; k++) {
                const int32_t t = h->h_et[h->h_re[k]];
                for (int64_t i = 0; i < h->h_nrows[t] - h->h_top[t]; i++) rowval[p++] = h->h_red_off[t] + i + 1;
            }
            colptr[c + 1] = p + 1;
        }
    return BLS_OK;
}

extern "C" int bls_check_reduced_jacobian_pattern(bls_qr *h, int64_t m, int64_t n, const int64_t *colptr,
                                                  const int64_t *rowval) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    const int64_t mm = h->red_len;
    if (m != mm || n != h->nr)
        return fail(BLS_ERR_ARG, "DimensionMismatch: Jt has size (" + std::to_string(m) + ", " + std::to_string(n) +
                                     "), expected (" + std::to_string(mm) + ", " + std::to_string(h->nr) + ")");
    if (!colptr || (colptr[n] > 1 && !rowval)) return fail(BLS_ERR_ARG, "bls_check_reduced_jacobian_pattern: bad arguments");
    for (int64_t r = 0; r < h->nrb; r++)
        for (int j = 0; j < h->h_r_dims[r]; j++) {
            const int64_t c = h->h_boff_r[r] + j;   // 0-based column
            if (colptr[c + 1] - colptr[c] != h->h_jh[r])
                return fail(BLS_ERR_ARG, "ArgumentError: Jt does not have the pattern of build_reduced_jacobian_pattern: "
                                         "column " + std::to_string(c + 1) + " stores " +
                                         std::to_string(colptr[c + 1] - colptr[c]) + " entries, expected " +
                                         std::to_string(h->h_jh[r]));
            int64_t p = colptr[c] - 1;
            for (int64_t k = h->h_retained_ptr[r]; k < h->h_retained_ptr[r + 1]; k++) {
                const int32_t t = h->h_et[h->h_re[k]];
                const int64_t nt = h->h_nrows[t] - h->h_top[t];
                if (nt && rowval[p] != h->h_red_off[t] + 1)
                    return fail(BLS_ERR_ARG, "ArgumentError: Jt does not have the pattern of "
                                             "build_reduced_jacobian_pattern: column " + std::to_string(c + 1) +
                                             " stores row " + std::to_string(rowval[p]) + ", expected " +
                                             std::to_string(h->h_red_off[t] + 1));
                p += nt;
            }
        }
    return BLS_OK;
}

namespace {
template <class T, bool SPARSE>
int reduced_jacobian(bls_qr *h, T *Jt, double *rt, const T *F, const double *Fr, hipStream_t s) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if ((!Jt && (SPARSE ? h->jac_nnz : h->red_len * h->nr)) || (h->red_len && (!rt || !Fr)) || (h->raw_len && !F))
        return fail(BLS_ERR_ARG, "reduced_jacobian: a required array is NULL");
    if (!SPARSE && h->red_len * h->nr)
        k_fill_zero<T><<<(unsigned)std::min<int64_t>(4096, grid_for(h->red_len * h->nr, 256)), 256, 0, s>>>(h->red_len * h->nr, Jt);
    if (h->ntables) {
        k_red_jacobian<T, SPARSE><<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), h->d_jdst.ptr, h->d_jh.ptr,
                                                                                  h->red_len, F, Jt);
        k_red_residual<<<(unsigned)((h->ntables + 3) / 4), 256, 0, s>>>(h->red(), Fr, rt);
    }
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
}  // namespace

extern "C" int bls_reduced_jacobian_sparse_f64(bls_qr *h, double *nzval, double *rt, const double *F, const double *Fr,
                                               void *stream) {
    return reduced_jacobian<double, true>(h, nzval, rt, F, Fr, (hipStream_t)stream);
}
extern "C" int bls_reduced_jacobian_sparse_f32(bls_qr *h, float *nzval, double *rt, const float *F, const double *Fr,
                                               void *stream) {
    return reduced_jacobian<float, true>(h, nzval, rt, F, Fr, (hipStream_t)stream);
}
extern "C" int bls_reduced_jacobian_dense_f64(bls_qr *h, double *Jt, double *rt, const double *F, const double *Fr,
                                              void *stream) {
    return reduced_jacobian<double, false>(h, Jt, rt, F, Fr, (hipStream_t)stream);
}
extern "C" int bls_reduced_jacobian_dense_f32(bls_qr *h, float *Jt, double *rt, const float *F, const double *Fr,
                                              void *stream) {
    return reduced_jacobian<float, false>(h, Jt, rt, F, Fr, (hipStream_t)stream);
}

extern "C" int bls_reduced_residual_f64(bls_qr *h, double *rt, const double *Fr, void *stream) {
    int rc = reduced_ready(h);
    if (rc) return rc;
    if (!rt || !Fr) return fail(BLS_ERR_ARG, "bls_reduced_residual_f64: a required array is NULL");
    if (h->ntables) k_red_residual<<<(unsigned)((h->ntables + 3) / 4), 256, 0, (hipStream_t)stream>>>(h->red(), Fr, rt);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_reduced_diagonal_blocks_f64(bls_qr *h, double *D, const double *F, void *stream) {
    return reduced_diag<double>(h, D, F, (hipStream_t)stream);
}
extern "C" int bls_reduced_diagonal_blocks_f32(bls_qr *h, float *D, const float *F, void *stream) {
    return reduced_diag<float>(h, D, F, (hipStream_t)stream);
}

extern "C" int bls_reduce_qr_f64(bls_qr *h, double *factored, double *fres, const double *raw, const double *r,
                                 const double *scaling_e, double lambda, const double *s, int64_t *failed_block,
                                 void *stream) {
    return reduce_qr_sync<double>(h, factored, fres, raw, r, scaling_e, lambda, s, failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_qr_f64_async(bls_qr *h, double *factored, double *fres, const double *raw, const double *r,
                                       const double *scaling_e, double lambda, const double *s, int64_t *d_failed_block,
                                       void *stream) {
    return reduce_qr_async<double>(h, factored, fres, raw, r, scaling_e, lambda, s, d_failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_qr_f32(bls_qr *h, float *factored, double *fres, const float *raw, const double *r,
                                 const double *scaling_e, double lambda, const double *s, int64_t *failed_block,
                                 void *stream) {
    return reduce_qr_sync<float>(h, factored, fres, raw, r, scaling_e, lambda, s, failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_qr_f32_async(bls_qr *h, float *factored, double *fres, const float *raw, const double *r,
                                       const double *scaling_e, double lambda, const double *s, int64_t *d_failed_block,
                                       void *stream) {
    return reduce_qr_async<float>(h, factored, fres, raw, r, scaling_e, lambda, s, d_failed_block, (hipStream_t)stream);
}

extern "C" int bls_qr_set_timing(bls_qr *h, int enable) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (enable) return h->timing.enable();
    h->timing.enabled = false;
    return BLS_OK;
}

extern "C" int bls_qr_kernel_time(bls_qr *h, float *ms) {
    if (!h || !ms) return fail(BLS_ERR_ARG, "bad arguments");
    *ms = h->timing.ms[0];
    return BLS_OK;
}
