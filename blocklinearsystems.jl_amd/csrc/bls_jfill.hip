// Filling the QR tables of a SchurBlockJacobian from per-row-block Jacobian blocks, on the device: the assignments
// block(J, k, blockentry(jst, k, j)) .= J_kj and residual_block(J, k) .= r_k (src/schur_jacobian.jl:330-387) for
// every row block k at once, behind the C ABI of include/bls_amd.h.
//
// The caller's values come row block after row block, each row block's blocks in ascending layout column block
// (the order of its sorted row_blocks), each m_k x d column-major; its residual rows likewise. Sixteen lanes take a
// row block (four per wavefront): lane q copies values q, q + 16, ... of the row block's run, so the reads are
// coalesced runs of 16 values, and writes each to its table column (stride M of the table). Row blocks own disjoint
// table rows, so the order of the writes does not matter.
#include <algorithm>

#include "bls_internal.h"

using namespace bls;

namespace {

struct JFRow {
    int64_t src;    // first value of the row block in the caller's array
    int64_t dst;    // raw offset of its table + its first row (0-based)
    int64_t rsrc;   // first residual row in the caller's array
    int64_t rdst;   // residual offset of its table + its first row
    int32_t M;      // rows of the table (column stride)
    int32_t m;      // rows of the row block
    int32_t b0;     // its first block in the block arrays
    int32_t nb;     // its blocks
};

constexpr int JF_LANES = 16;

template <class T>
__global__ __launch_bounds__(256) void k_jfill(int64_t nrows, const JFRow *rows, const int32_t *bcol,
                                               const int32_t *bdim, const T *vals, const double *res, T *raw,
                                               double *r) {
    const int64_t k = (int64_t)blockIdx.x * (256 / JF_LANES) + threadIdx.x / JF_LANES;
    if (k >= nrows) return;
    const int lane = threadIdx.x % JF_LANES;
    const JFRow rb = rows[k];
    int W = 0;
    for (int j = 0; j < rb.nb; j++) W += bdim[rb.b0 + j];
    const int64_t n = (int64_t)rb.m * W;
    for (int64_t v = lane; v < n; v += JF_LANES) {
        const int cc = (int)(v / rb.m), i = (int)(v % rb.m);
        int j = 0, pre = 0;
        while (cc >= pre + bdim[rb.b0 + j]) pre += bdim[rb.b0 + j++];
        const int64_t col = bcol[rb.b0 + j] + (cc - pre);
        raw[rb.dst + col * rb.M + i] = vals[rb.src + v];
    }
    for (int i = lane; i < rb.m; i += JF_LANES) r[rb.rdst + i] = res[rb.rsrc + i];
}

}  // namespace

struct bls_jfill {
    int64_t nrowblocks = 0, nvals = 0, nres = 0, raw_len = 0, res_len = 0;
    DevArray<JFRow> d_rows;
    DevArray<int32_t> d_bcol, d_bdim;
};

extern "C" int bls_jfill_create(int64_t nrblocks, int64_t neblocks, const int64_t *r_dims, const int64_t *e_dims,
                                int64_t ntables, const int64_t *table_nrows, const int64_t *table_ncols,
                                const int64_t *table_top, const int64_t *raw_offsets, const int64_t *residual_offsets,
                                const int64_t *table_ptr, const int64_t *entry_retained, const int64_t *entry_col,
                                int64_t nrowblocks, const int64_t *row_table, const int64_t *row_first,
                                const int64_t *row_nrows, const int64_t *row_ptr, const int64_t *row_blocks,
                                int64_t *nvals, int64_t *nres, bls_jfill **out) {
    if (!out || nrblocks < 0 || neblocks < 0 || ntables < neblocks || nrowblocks < 0 || !table_nrows ||
        !table_ncols || !table_top || !raw_offsets || !residual_offsets || !table_ptr || !row_ptr ||
        (nrblocks && !r_dims) || (neblocks && !e_dims) ||
        (nrowblocks && (!row_table || !row_first || !row_nrows || !row_blocks)))
        return fail(BLS_ERR_ARG, "bls_jfill_create: bad arguments");
    *out = nullptr;
    std::vector<JFRow> rows((size_t)nrowblocks);
    std::vector<int32_t> bcol, bdim;
    int64_t src = 0, rsrc = 0;
    for (int64_t k = 0; k < nrowblocks; k++) {
        const int64_t t = row_table[k] - 1;
        if (t < 0 || t >= ntables) return fail(BLS_ERR_ARG, "row block " + std::to_string(k + 1) + ": table out of range");
        const int64_t M = table_nrows[t];
        const int64_t data_rows = M - (t < neblocks ? table_top[t] : 0);   // data_rows (src/schur_jacobian.jl:251)
        const int64_t first = row_first[k] - 1, m = row_nrows[k];
        if (m < 0 || first < 0 || first + m > data_rows)
            return fail(BLS_ERR_ARG, "row block " + std::to_string(k + 1) + ": rows outside the data rows of table " +
                                         std::to_string(t + 1));
        JFRow rb{};
        rb.src = src;
        rb.dst = raw_offsets[t] - 1 + first;
        rb.rsrc = rsrc;
        rb.rdst = residual_offsets[t] - 1 + first;
        rb.M = (int32_t)M;
        rb.m = (int32_t)m;
        rb.b0 = (int32_t)bcol.size();
        int64_t W = 0;
        for (int64_t q = row_ptr[k] - 1; q < row_ptr[k + 1] - 1; q++) {
            const int64_t j = row_blocks[q];   // layout column block: retained 1..nrblocks, eliminated nrblocks + e
            int64_t col, d;
            if (j >= 1 && j <= nrblocks) {
                // coupling_entry (src/schur_jacobian.jl:181-187): the entry of retained block j in table t
                const int64_t *lo = entry_retained + (table_ptr[t] - 1), *hi = entry_retained + (table_ptr[t + 1] - 1);
                const int64_t *it = std::lower_bound(lo, hi, j);
                if (it == hi || *it != j)
                    return fail(BLS_ERR_ARG, "Row block " + std::to_string(k + 1) + " does not touch column block " +
                                                 std::to_string(j));
                const int64_t u = it - entry_retained;
                col = entry_col[u] - 1;
                d = r_dims[j - 1];
            } else if (j > nrblocks && j - nrblocks == t + 1 && t < neblocks) {
                d = e_dims[t];
                col = table_ncols[t] - 1 - d;   // eliminated_block: columns table_ncols - d_e, ... (1-based)
            } else {
                return fail(BLS_ERR_ARG, "Row block " + std::to_string(k + 1) + " does not touch column block " +
                                             std::to_string(j));
            }
            bcol.push_back((int32_t)col);
            bdim.push_back((int32_t)d);
            W += d;
        }
        rb.nb = (int32_t)(bcol.size() - rb.b0);
        rows[k] = rb;
        src += m * W;
        rsrc += m;
    }
    auto *h = new bls_jfill();
    h->nrowblocks = nrowblocks;
    h->nvals = src;
    h->nres = rsrc;
    h->raw_len = raw_offsets[ntables] - 1;
    h->res_len = residual_offsets[ntables] - 1;
    int rc;
    if ((rc = h->d_rows.upload(rows)) || (rc = h->d_bcol.upload(bcol)) || (rc = h->d_bdim.upload(bdim))) {
        delete h;
        return rc;
    }
    if (nvals) *nvals = src;
    if (nres) *nres = rsrc;
    *out = h;
    return BLS_OK;
}

extern "C" void bls_jfill_destroy(bls_jfill *h) { delete h; }

namespace {
template <class T> int jfill(bls_jfill *h, T *raw, double *r, const T *vals, const double *res, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if ((h->nvals && (!raw || !vals)) || (h->nres && (!r || !res)))
        return fail(BLS_ERR_ARG, "jfill: a required array is NULL");
    if (h->nrowblocks)
        k_jfill<T><<<grid_for(h->nrowblocks, 256 / JF_LANES), 256, 0, s>>>(h->nrowblocks, h->d_rows.ptr, h->d_bcol.ptr,
                                                                          h->d_bdim.ptr, vals, res, raw, r);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
}  // namespace

extern "C" int bls_jfill_f64(bls_jfill *h, double *raw, double *r, const double *vals, const double *res,
                             void *stream) {
    return jfill<double>(h, raw, r, vals, res, (hipStream_t)stream);
}

extern "C" int bls_jfill_f32(bls_jfill *h, float *raw, double *r, const float *vals, const double *res,
                             void *stream) {
    return jfill<float>(h, raw, r, vals, res, (hipStream_t)stream);
}
