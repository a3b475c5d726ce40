/*
 * CPU oracle for the sparse-solver path — TEST INFRASTRUCTURE ONLY.
 *
 * Plain C restatement of the reference algorithms, used by tests/, the
 * package smoke check and the bench's cpu_baseline leg as a checker and a
 * serial CPU baseline.  The product path never links or calls this file.
 *
 * Every function cites the reference code it restates (paths relative to the
 * reference checkout, NVIDIA Warp 1.17.0.dev4).
 *
 * Built with `-O2 -ffp-contract=fast -mfma` to mirror the reference CPU JIT
 * flags (warp/_src/build.py:154-200, warp/native/clang/clang.cpp:214-217).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Triplets -> compact BSR                                                    */
/* ------------------------------------------------------------------------- */

/* Zero test of one block under a scalar bit mask.
 * Restates warp/native/sparse.cpp:22-29 (bsr_block_is_zero) and
 * warp/native/sparse.cu:245-268 (BsrBlockIsNotZero). */
static int orc_block_is_zero(const void* values, int scalar_bytes, int block_size, int64_t idx, uint64_t mask)
{
    const unsigned char* base = (const unsigned char*)values + (size_t)idx * block_size * scalar_bytes;
    for (int k = 0; k < block_size; ++k) {
        uint64_t bits = 0;
        memcpy(&bits, base + (size_t)k * scalar_bytes, scalar_bytes);
        if ((bits & mask) != 0)
            return 0;
    }
    return 1;
}

typedef struct {
    const int* rows;
    const int* cols;
} orc_key_ctx;

static int orc_key_less(const orc_key_ctx* c, int i, int j)
{
    return c->rows[i] < c->rows[j] || (c->rows[i] == c->rows[j] && c->cols[i] < c->cols[j]);
}

/* Stable bottom-up merge sort of triplet indices by (row, col).
 * The reference CUDA path sorts (row<<32|col) keys with a stable LSD radix
 * sort carrying the triplet index (warp/native/sparse.cu:1250-1263), so equal
 * keys keep ascending triplet order; this merge sort reproduces that order. */
static void orc_stable_sort(int* idx, int n, const orc_key_ctx* c)
{
    if (n < 2)
        return;
    int* tmp = (int*)malloc(sizeof(int) * (size_t)n);
    int* src = idx;
    int* dst = tmp;
    for (int width = 1; width < n; width *= 2) {
        for (int lo = 0; lo < n; lo += 2 * width) {
            int mid = lo + width < n ? lo + width : n;
            int hi = lo + 2 * width < n ? lo + 2 * width : n;
            int a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                /* take from the right run only when strictly smaller: stable */
                if (orc_key_less(c, src[b], src[a]))
                    dst[o++] = src[b++];
                else
                    dst[o++] = src[a++];
            }
            while (a < mid)
                dst[o++] = src[a++];
            while (b < hi)
                dst[o++] = src[b++];
        }
        int* t = src;
        src = dst;
        dst = t;
    }
    if (src != idx)
        memcpy(idx, src, sizeof(int) * (size_t)n);
    free(tmp);
}

/*
 * Compact BSR topology (+ summed values) from COO triplets.
 *
 * Topology: warp/native/sparse.cpp:267-413 (host) / sparse.cu:1168-1322
 * (device): drop out-of-range triplets and, when `zero_mask` != 0 and values
 * are given, blocks whose scalars are all zero under the mask; sort by
 * (row, col) stably; one output block per distinct key; offsets from per-row
 * counts.  `summed_offsets[u]` is the inclusive end of run u in the sorted
 * order, `summed_indices` the sorted triplet indices.
 *
 * Values: warp/_src/sparse.py:1122-1141 (_bsr_accumulate_triplet_value):
 * values[u] = v[i0]; values[u] += v[i1]; ... in ascending triplet order, in
 * the block scalar type (scalar_bytes 4 = float, 8 = double).
 *
 * `count` (nullable) limits the triplets to the first count[0]
 * (sparse.cu:287-288, sparse.cpp:291-293).  Returns nnz = offsets[nrow].
 * summed_offsets / summed_indices / columns must hold T entries.
 */
ORC_API int orc_bsr_from_triplets(
    int block_size, int scalar_bytes, int nrow, int ncol, int T, const int* count, const int* rows, const int* cols,
    const void* values, uint64_t zero_mask, int* summed_offsets, int* summed_indices, int* offsets, int* columns,
    void* out_values
)
{
    int n = T;
    if (count)
        n = *count < T ? *count : T;
    if (n < 0)
        n = 0;

    int valid = 0;
    for (int i = 0; i < n; ++i) {
        int r = rows[i], c = cols[i];
        if (r < 0 || r >= nrow || c < 0 || c >= ncol)
            continue;
        if (values && zero_mask && orc_block_is_zero(values, scalar_bytes, block_size, i, zero_mask))
            continue;
        summed_indices[valid++] = i;
    }

    orc_key_ctx ctx = {rows, cols};
    orc_stable_sort(summed_indices, valid, &ctx);

    for (int r = 0; r <= nrow; ++r)
        offsets[r] = 0;

    int nnz = 0;
    int cur_r = -1, cur_c = -1;
    for (int k = 0; k < valid; ++k) {
        int i = summed_indices[k];
        int r = rows[i], c = cols[i];
        if (r != cur_r || c != cur_c) {
            columns[nnz] = c;
            offsets[r + 1] += 1;
            ++nnz;
            cur_r = r;
            cur_c = c;
        }
        summed_offsets[nnz - 1] = k + 1;
    }
    for (int r = 0; r < nrow; ++r)
        offsets[r + 1] += offsets[r];

    if (values && out_values) {
        for (int u = 0; u < nnz; ++u) {
            int beg = u == 0 ? 0 : summed_offsets[u - 1];
            int end = summed_offsets[u];
            for (int s = 0; s < block_size; ++s) {
                if (scalar_bytes == 4) {
                    const float* v = (const float*)values;
                    float acc = v[(size_t)summed_indices[beg] * block_size + s];
                    for (int k = beg + 1; k < end; ++k)
                        acc += v[(size_t)summed_indices[k] * block_size + s];
                    ((float*)out_values)[(size_t)u * block_size + s] = acc;
                } else {
                    const double* v = (const double*)values;
                    double acc = v[(size_t)summed_indices[beg] * block_size + s];
                    for (int k = beg + 1; k < end; ++k)
                        acc += v[(size_t)summed_indices[k] * block_size + s];
                    ((double*)out_values)[(size_t)u * block_size + s] = acc;
                }
            }
        }
    }
    return nnz;
}

/* ------------------------------------------------------------------------- */
/* SpMV                                                                       */
/* ------------------------------------------------------------------------- */

/* y = alpha * A x + beta * y, reference row kernel
 * warp/_src/sparse.py:4627-4665 (bsr_mv_kernel): per scalar row, sum over
 * blocks in storage order and block columns in order, then scale by alpha,
 * then add beta*y.  alpha == 0 leaves x unread, beta == 0 leaves y unread. */
#define ORC_DEFINE_BSR_MV(SUFFIX, S)                                                                                  \
    ORC_API void orc_bsr_mv_##SUFFIX(                                                                                 \
        int nrow, int br, int bc, const int* offsets, const int* row_counts, const int* columns, const S* values,     \
        const S* x, S* y, S alpha, S beta                                                                             \
    )                                                                                                                 \
    {                                                                                                                 \
        for (int row = 0; row < nrow; ++row) {                                                                        \
            int beg = offsets[row];                                                                                   \
            int end = row_counts ? beg + row_counts[row] : offsets[row + 1];                                          \
            for (int s = 0; s < br; ++s) {                                                                            \
                S v = (S)0;                                                                                           \
                if (alpha != (S)0) {                                                                                  \
                    for (int b = beg; b < end; ++b) {                                                                 \
                        const S* blk = values + (size_t)b * br * bc + (size_t)s * bc;                                 \
                        const S* xs = x + (size_t)columns[b] * bc;                                                    \
                        for (int c = 0; c < bc; ++c)                                                                  \
                            v += blk[c] * xs[c];                                                                      \
                    }                                                                                                 \
                    v *= alpha;                                                                                       \
                }                                                                                                     \
                if (beta != (S)0)                                                                                     \
                    v += beta * y[(size_t)row * br + s];                                                              \
                y[(size_t)row * br + s] = v;                                                                          \
            }                                                                                                         \
        }                                                                                                             \
    }

ORC_DEFINE_BSR_MV(f32, float)
ORC_DEFINE_BSR_MV(f64, double)

/* Same product in the lane-group order of the GPU vector-BSR kernels
 * (warp_amd/csrc/bsr_core.h): for G lanes per block row, lane g sums blocks
 * g, g+G, ... (block columns in order, fused multiply-adds), then lanes are
 * combined by an xor butterfly (off = G/2 .. 1); the result is scaled by alpha
 * and beta*y is added with one fused multiply-add.  G == 0 falls back to the
 * reference row order above.  Reproduces the GPU bits exactly; it does not
 * restate the reference (whose FP32 order is unpinnable). */
#define ORC_DEFINE_BSR_MV_GROUPED(SUFFIX, S, FMA)                                                                     \
    ORC_API void orc_bsr_mv_grouped_##SUFFIX(                                                                         \
        int nrow, int br, int bc, const int* offsets, const int* row_counts, const int* columns, const S* values,     \
        const S* x, S* y, S alpha, S beta, int G                                                                      \
    )                                                                                                                 \
    {                                                                                                                 \
        if (G <= 0) {                                                                                                 \
            orc_bsr_mv_##SUFFIX(nrow, br, bc, offsets, row_counts, columns, values, x, y, alpha, beta);              \
            return;                                                                                                   \
        }                                                                                                             \
        S a[64], t[64];                                                                                               \
        for (int row = 0; row < nrow; ++row) {                                                                        \
            int beg = offsets[row];                                                                                   \
            int end = row_counts ? beg + row_counts[row] : offsets[row + 1];                                          \
            for (int s = 0; s < br; ++s) {                                                                            \
                for (int g = 0; g < G; ++g) {                                                                         \
                    a[g] = (S)0;                                                                                      \
                    if (alpha == (S)0)                                                                                \
                        continue;                                                                                     \
                    for (int b = beg + g; b < end; b += G) {                                                          \
                        const S* blk = values + (size_t)b * br * bc + (size_t)s * bc;                                 \
                        const S* xs = x + (size_t)columns[b] * bc;                                                    \
                        for (int c = 0; c < bc; ++c)                                                                  \
                            a[g] = FMA(blk[c], xs[c], a[g]);                                                          \
                    }                                                                                                 \
                }                                                                                                     \
                for (int off = G / 2; off > 0; off >>= 1) {                                                           \
                    for (int g = 0; g < G; ++g)                                                                       \
                        t[g] = a[g] + a[g ^ off];                                                                     \
                    for (int g = 0; g < G; ++g)                                                                       \
                        a[g] = t[g];                                                                                  \
                }                                                                                                     \
                S v = a[0] * alpha;                                                                                   \
                if (beta != (S)0)                                                                                     \
                    v = FMA(beta, y[(size_t)row * br + s], v);                                                        \
                y[(size_t)row * br + s] = v;                                                                          \
            }                                                                                                         \
        }                                                                                                             \
    }

ORC_DEFINE_BSR_MV_GROUPED(f32, float, fmaf)
ORC_DEFINE_BSR_MV_GROUPED(f64, double, fma)

/* ------------------------------------------------------------------------- */
/* Diagonal and Jacobi                                                        */
/* ------------------------------------------------------------------------- */

/* Binary search of column `col` among the sorted active blocks of `row`.
 * Restates warp/_src/sparse.py:1852-1905 (bsr_block_index). */
static int orc_block_index(int row, int col, const int* offsets, const int* row_counts, const int* columns)
{
    int lo = offsets[row];
    int hi = row_counts ? lo + row_counts[row] : offsets[row + 1];
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (columns[mid] < col)
            lo = mid + 1;
        else
            hi = mid;
    }
    int end = row_counts ? offsets[row] + row_counts[row] : offsets[row + 1];
    return (lo < end && columns[lo] == col) ? lo : -1;
}

/* Diagonal blocks (zero where absent): warp/_src/sparse.py:2871-2924.
 * Inverse diagonal coefficients: warp/_src/optim/linear.py:2290-2320
 * (1/d, or 1/|d| with use_abs, and 1 where d == 0); for a block matrix the
 * inverse of the diagonal *entries* of each diagonal block (point Jacobi). */
#define ORC_DEFINE_DIAG(SUFFIX, S)                                                                                    \
    ORC_API void orc_bsr_get_diag_##SUFFIX(                                                                           \
        int nrow, int ncol, int br, int bc, const int* offsets, const int* row_counts, const int* columns,            \
        const S* values, S* out                                                                                       \
    )                                                                                                                 \
    {                                                                                                                 \
        int dim = nrow < ncol ? nrow : ncol;                                                                          \
        for (int r = 0; r < dim; ++r) {                                                                               \
            int b = orc_block_index(r, r, offsets, row_counts, columns);                                              \
            for (int k = 0; k < br * bc; ++k)                                                                         \
                out[(size_t)r * br * bc + k] = b >= 0 ? values[(size_t)b * br * bc + k] : (S)0;                       \
        }                                                                                                             \
    }                                                                                                                 \
    ORC_API void orc_inverse_diag_##SUFFIX(                                                                           \
        int dim, int br, int bc, const S* diag_blocks, int use_abs, S* inv_diag                                       \
    )                                                                                                                 \
    {                                                                                                                 \
        for (int r = 0; r < dim; ++r)                                                                                 \
            for (int s = 0; s < br; ++s) {                                                                            \
                S d = s < bc ? diag_blocks[(size_t)r * br * bc + (size_t)s * bc + s] : (S)0;                          \
                S den = use_abs ? (d < 0 ? -d : d) : d;                                                               \
                inv_diag[(size_t)r * br + s] = d == (S)0 ? (S)1 : (S)1 / den;                                         \
            }                                                                                                         \
    }

ORC_DEFINE_DIAG(f32, float)
ORC_DEFINE_DIAG(f64, double)

/* ------------------------------------------------------------------------- */
/* Conjugate gradient                                                         */
/* ------------------------------------------------------------------------- */

/* Preconditioner kinds for the oracle CG. */
enum { ORC_PREC_NONE = 0, ORC_PREC_DIAG = 1, ORC_PREC_BLOCK = 2 };

/* Tolerance defaults: warp/_src/optim/linear.py:2010-2032.  `tol`/`atol`
 * NaN stands for None.  eps = 1e-16 (fp64) or 1e-8 (other). */
static void orc_get_tolerances(int is_f64, double tol, double atol, double* rtol_out, double* atol_out)
{
    double eps = is_f64 ? 1.0e-16 : 1.0e-8;
    double default_tol = pow(eps, 0.75);
    double min_tol = pow(eps, 2.25);
    if (isnan(tol) && isnan(atol)) {
        tol = atol = default_tol;
    } else if (isnan(tol)) {
        tol = atol;
    } else if (isnan(atol)) {
        atol = tol;
    }
    if (atol < min_tol)
        atol = min_tol;
    *rtol_out = tol;
    *atol_out = atol;
}

/*
 * Preconditioned CG with the reference loop semantics:
 *   warp/_src/optim/linear.py:893-1033 (CG._allocate/_run),
 *   2051-2083 (absolute tolerance from ||b||), 2182-2243 (host loop: first
 *   check at iteration 0, then every `check_every` iterations, maxiter cut),
 *   2332-2373 (_cg_kernel_1/2: alpha and beta frozen to 0 once r.r <= atol^2).
 * `restart` > 0 enables the restart cycle (linear.py:1005-1021, 872-876).
 *
 * Vectors are in S; dot products and alpha/beta in D.  With D == S this is
 * the reference's precision; with D = double it is the FP64-scalar policy of
 * the GPU solver.  Dots are summed in 512-wide tiles, then over tiles
 * (the tile structure of linear.py:309-612; order inside a tile is
 * sequential here).
 *
 * Operator: BSR (br == bc); `group` 0 sums rows in the reference order
 * (orc_bsr_mv), G > 0 in the GPU lane-group order (orc_bsr_mv_grouped);
 * vector updates are single fused multiply-adds.  Preconditioner: none,
 * diagonal (`prec` holds n scalars), or block (`prec` holds nrow br x br
 * blocks, row-major, applied as z = P r per block row).
 *
 * out[0] = iterations, out[1] = ||r|| (recursive residual), out[2] = atol.
 * `history` (nullable, capacity `history_cap`) receives ||r|| at every
 * check, as the reference callback would see it; out[3] = entries written.
 */
#define ORC_DEFINE_CG(SUFFIX, S, D, IS_F64, FMA)                                                                           \
    static D orc_dot_##SUFFIX(const S* a, const S* b, int n)                                                          \
    {                                                                                                                 \
        D total = (D)0;                                                                                               \
        for (int t = 0; t < n; t += 512) {                                                                            \
            int e = t + 512 < n ? t + 512 : n;                                                                        \
            D part = (D)0;                                                                                            \
            for (int i = t; i < e; ++i)                                                                               \
                part += (D)a[i] * (D)b[i];                                                                            \
            total += part;                                                                                            \
        }                                                                                                             \
        return total;                                                                                                 \
    }                                                                                                                 \
    static void orc_prec_apply_##SUFFIX(                                                                              \
        int kind, int nrow, int br, const S* prec, const S* r, S* z                                                   \
    )                                                                                                                 \
    {                                                                                                                 \
        int n = nrow * br;                                                                                            \
        if (kind == ORC_PREC_DIAG) {                                                                                  \
            for (int i = 0; i < n; ++i)                                                                               \
                z[i] = prec[i] * r[i];                                                                                \
        } else if (kind == ORC_PREC_BLOCK) {                                                                          \
            for (int b = 0; b < nrow; ++b)                                                                            \
                for (int s = 0; s < br; ++s) {                                                                        \
                    S v = (S)0;                                                                                       \
                    for (int c = 0; c < br; ++c)                                                                      \
                        v = FMA(prec[(size_t)b * br * br + (size_t)s * br + c], r[(size_t)b * br + c], v);            \
                    z[(size_t)b * br + s] = v;                                                                        \
                }                                                                                                     \
        } else {                                                                                                      \
            memcpy(z, r, sizeof(S) * (size_t)n);                                                                      \
        }                                                                                                             \
    }                                                                                                                 \
    ORC_API void orc_cg_##SUFFIX(                                                                                     \
        int nrow, int br, const int* offsets, const int* columns, const S* values, const S* b, S* x, double tol,      \
        double atol, int maxiter, int prec_kind, const S* prec, int check_every, int restart, int group, double* out, \
        double* history, int history_cap                                                                              \
    )                                                                                                                 \
    {                                                                                                                 \
        int n = nrow * br;                                                                                            \
        S* r = (S*)calloc((size_t)n, sizeof(S));                                                                      \
        S* z = (S*)calloc((size_t)n, sizeof(S));                                                                      \
        S* p = (S*)calloc((size_t)n, sizeof(S));                                                                      \
        S* Ap = (S*)calloc((size_t)n, sizeof(S));                                                                     \
        int nhist = 0;                                                                                                \
        if (maxiter <= 0)                                                                                             \
            maxiter = n;                                                                                              \
        int cycle = 1;                                                                                                \
        if (restart > 0) {                                                                                            \
            cycle = restart < maxiter ? restart : maxiter;                                                            \
            if (cycle < 1)                                                                                            \
                cycle = 1;                                                                                            \
            if (check_every > 0 && check_every < cycle)                                                               \
                check_every = cycle;                                                                                  \
        }                                                                                                             \
        double rtol_d, atol_d;                                                                                        \
        orc_get_tolerances(IS_F64, tol, atol, &rtol_d, &atol_d);                                                      \
        D bb = orc_dot_##SUFFIX(b, b, n);                                                                             \
        D a_tol = (D)rtol_d * (D)sqrt((double)bb);                                                                    \
        if (a_tol < (D)atol_d)                                                                                        \
            a_tol = (D)atol_d;                                                                                        \
        D atol_sq = a_tol * a_tol;                                                                                    \
        /* r = b - A x  (LinearOperator.matvec(x, b, r, -1, 1), linear.py:159-164) */                                  \
        memcpy(r, b, sizeof(S) * (size_t)n);                                                                          \
        orc_bsr_mv_##SUFFIX##_op(nrow, br, offsets, columns, values, x, r, (S)-1, (S)1, group);                              \
        D rr, rz;                                                                                                     \
        orc_prec_apply_##SUFFIX(prec_kind, nrow, br, prec, r, z);                                                     \
        rr = orc_dot_##SUFFIX(r, r, n);                                                                               \
        rz = prec_kind ? orc_dot_##SUFFIX(r, z, n) : rr;                                                              \
        memcpy(p, z, sizeof(S) * (size_t)n);                                                                          \
        int it = 0;                                                                                                   \
        if (history && nhist < history_cap)                                                                           \
            history[nhist++] = sqrt((double)rr);                                                                      \
        if (!(rr <= atol_sq)) {                                                                                       \
            for (;;) {                                                                                                \
                for (int k = 0; k < cycle; ++k) {                                                                     \
                    int update_direction = !(restart > 0 && k == cycle - 1);                                          \
                    D rz_old = rz;                                                                                    \
                    orc_bsr_mv_##SUFFIX##_op(nrow, br, offsets, columns, values, p, Ap, (S)1, (S)0, group);                  \
                    D pAp = orc_dot_##SUFFIX(p, Ap, n);                                                               \
                    D alpha = rr > atol_sq ? rz_old / pAp : (D)0;                                                     \
                    for (int i = 0; i < n; ++i) {                                                                     \
                        x[i] = FMA((S)alpha, p[i], x[i]);                                                             \
                        r[i] = FMA(-(S)alpha, Ap[i], r[i]);                                                           \
                    }                                                                                                 \
                    if (update_direction) {                                                                           \
                        orc_prec_apply_##SUFFIX(prec_kind, nrow, br, prec, r, z);                                     \
                        rr = orc_dot_##SUFFIX(r, r, n);                                                               \
                        rz = prec_kind ? orc_dot_##SUFFIX(r, z, n) : rr;                                              \
                        D beta = rr > atol_sq ? rz / rz_old : (D)0;                                                   \
                        const S* zz = prec_kind ? z : r;                                                              \
                        for (int i = 0; i < n; ++i)                                                                   \
                            p[i] = FMA((S)beta, p[i], zz[i]);                                                         \
                    }                                                                                                 \
                }                                                                                                     \
                if (restart > 0) {                                                                                    \
                    memcpy(r, b, sizeof(S) * (size_t)n);                                                              \
                    orc_bsr_mv_##SUFFIX##_op(nrow, br, offsets, columns, values, x, r, (S)-1, (S)1, group);                  \
                    orc_prec_apply_##SUFFIX(prec_kind, nrow, br, prec, r, z);                                         \
                    rr = orc_dot_##SUFFIX(r, r, n);                                                                   \
                    rz = prec_kind ? orc_dot_##SUFFIX(r, z, n) : rr;                                                  \
                    memcpy(p, prec_kind ? z : r, sizeof(S) * (size_t)n);                                              \
                }                                                                                                     \
                it += cycle;                                                                                          \
                if (it >= maxiter)                                                                                    \
                    break;                                                                                            \
                if (check_every == 0) {                                                                               \
                    /* device-side condition after every cycle (linear.py:2096-2116) */                              \
                    if (rr <= atol_sq)                                                                                \
                        break;                                                                                        \
                } else if ((it % check_every) < cycle) {                                                              \
                    if (rr <= atol_sq)                                                                                \
                        break;                                                                                        \
                    if (history && nhist < history_cap)                                                               \
                        history[nhist++] = sqrt((double)rr);                                                          \
                }                                                                                                     \
            }                                                                                                         \
            if (history && nhist < history_cap)                                                                       \
                history[nhist++] = sqrt((double)rr);                                                                  \
        }                                                                                                             \
        out[0] = (double)it;                                                                                          \
        out[1] = sqrt((double)rr);                                                                                    \
        out[2] = sqrt((double)atol_sq);                                                                               \
        out[3] = (double)nhist;                                                                                       \
        free(r);                                                                                                      \
        free(z);                                                                                                      \
        free(p);                                                                                                      \
        free(Ap);                                                                                                     \
    }

/* Square-block operator adaptor for the CG loop (group 0 = reference order). */
static void orc_bsr_mv_f32_op(
    int nrow, int br, const int* offsets, const int* columns, const float* values, const float* x, float* y,
    float alpha, float beta, int group
)
{
    orc_bsr_mv_grouped_f32(nrow, br, br, offsets, NULL, columns, values, x, y, alpha, beta, group);
}
static void orc_bsr_mv_f64_op(
    int nrow, int br, const int* offsets, const int* columns, const double* values, const double* x, double* y,
    double alpha, double beta, int group
)
{
    orc_bsr_mv_grouped_f64(nrow, br, br, offsets, NULL, columns, values, x, y, alpha, beta, group);
}
#define orc_bsr_mv_f32d_op orc_bsr_mv_f32_op

ORC_DEFINE_CG(f32, float, float, 0, fmaf)
ORC_DEFINE_CG(f64, double, double, 1, fma)
ORC_DEFINE_CG(f32d, float, double, 0, fmaf)

/* ------------------------------------------------------------------------- */
/* FP64 residual for mixed precision                                          */
/* ------------------------------------------------------------------------- */

/* r = b - A x with FP32 matrix values promoted to FP64 and FP64 vectors.
 * No reference counterpart (iterative refinement is new on this path); the
 * summation order is the row order of orc_bsr_mv. */
ORC_API void orc_residual_f32_f64(
    int nrow, int br, const int* offsets, const int* columns, const float* values, const double* x, const double* b,
    double* r
)
{
    for (int row = 0; row < nrow; ++row)
        for (int s = 0; s < br; ++s) {
            double v = 0.0;
            for (int blk = offsets[row]; blk < offsets[row + 1]; ++blk)
                for (int c = 0; c < br; ++c)
                    v += (double)values[(size_t)blk * br * br + (size_t)s * br + c] * x[(size_t)columns[blk] * br + c];
            r[(size_t)row * br + s] = b[(size_t)row * br + s] - v;
        }
}
