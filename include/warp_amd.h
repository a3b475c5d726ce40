/*
 * warp_amd.h -- C ABI of libwarp_amd.so, the gfx950 (MI355X) library behind
 * the sparse-solver path of Warp (bsr_set_from_triplets, bsr_mv, Jacobi,
 * conjugate gradient).
 *
 * Conventions
 *   - Every pointer is caller-owned device memory unless documented as host.
 *   - `stream` is a hipStream_t passed as void* (NULL = legacy default stream);
 *     all work is stream-ordered and asynchronous unless an entry point
 *     documents a host read-back.
 *   - int-returning entry points return 0 on success, a negative status on a
 *     HIP or argument error; warp_amd_last_error() then describes it (per
 *     host thread).
 *   - Entry points taking (workspace, workspace_bytes) never allocate; the
 *     matching *_workspace_size() query gives the byte count (host-only
 *     call, no GPU needed).  A workspace is 16-byte aligned (device
 *     allocations are); the CG solvers reject one that is not.
 *   - Reference paths below are relative to the NVIDIA Warp 1.17.0.dev4
 *     checkout.
 */
#ifndef WARP_AMD_H
#define WARP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#endif

/* ------------------------------------------------------------------------ */
/* Runtime                                                                  */
/* ------------------------------------------------------------------------ */

/* Message of the last failure on this host thread
 * (replaces wp_get_error_string, warp/native/warp.h:38). */
const char* warp_amd_last_error(void);
void warp_amd_clear_error(void);
/* major * 10000 + minor * 100 + patch */
int warp_amd_version(void);
int warp_amd_device_count(int* count);

/* Memory helpers for callers without their own allocator
 * (replace wp_alloc_device / wp_free_device / wp_memcpy_*, warp/native/warp.h,
 * implemented at warp/native/warp.cu:752-866, 1087).
 * memcpy kind: 0 host->device, 1 device->host, 2 device->device. */
int warp_amd_malloc(void** ptr, size_t bytes);
int warp_amd_free(void* ptr);
int warp_amd_malloc_host(void** ptr, size_t bytes);
int warp_amd_free_host(void* ptr);
int warp_amd_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int warp_amd_stream_synchronize(void* stream);

/* Per-thread "current stream" used by the reference-named entry point below
 * (the reference uses its current CUDA stream, warp/native/sparse.cu:1189-1192). */
void warp_amd_set_current_stream(void* stream);
void* warp_amd_get_current_stream(void);

/* ------------------------------------------------------------------------ */
/* Triplets -> compact BSR                                                  */
/* ------------------------------------------------------------------------ */

size_t warp_amd_bsr_from_triplets_workspace_size(int row_count, int nnz);

/* Replaces wp_bsr_matrix_from_triplets_device (warp/native/warp.h:563-602,
 * warp/native/sparse.cu:1168-1322) fused with the value accumulation kernel
 * _bsr_accumulate_triplet_values (warp/_src/sparse.py:1073-1090, 1122-1141).
 *
 * nnz            number of triplets T (upper bound; tpl_nnz, when non-null,
 *                is a device int32 holding the actual count <= T)
 * tpl_values     T blocks of block_size scalars (scalar_size bytes each), or
 *                NULL for a topology-only build
 * scalar_zero_mask  non-zero: drop blocks whose scalars all satisfy
 *                (bits & mask) == 0 (sparse.py:993-1006)
 * masked_topology   1: keep only triplets present in the existing topology
 *                (bsr_offsets / bsr_row_counts / bsr_columns are read first)
 * summed_block_offsets / _indices  optional int32[T] outputs: inclusive run
 *                end of each unique block in sorted order, and the triplet
 *                indices in (row, col, triplet) order; NULL = kept internal
 * bsr_offsets    int32[row_count + 1] output (input too when masked)
 * bsr_columns    int32[>= unique count] output
 * bsr_values     NULL, or [>= unique count] blocks receiving the sums
 *                v[i0] + v[i1] + ... in ascending triplet order (fp32 / fp64)
 * bsr_nnz_host   NULL, or pinned host int32 receiving offsets[row_count]
 *                asynchronously (valid once the stream reaches this point)
 * workspace      >= warp_amd_bsr_from_triplets_workspace_size(row_count, nnz)
 */
int warp_amd_bsr_from_triplets(
    int block_size, int scalar_size, int row_count, int col_count, int nnz, const int* tpl_nnz,
    const int* tpl_rows, const int* tpl_columns, const void* tpl_values, uint64_t scalar_zero_mask,
    int masked_topology, int* summed_block_offsets, int* summed_block_indices, int* bsr_offsets,
    const int* bsr_row_counts, int* bsr_columns, void* bsr_values, int* bsr_nnz_host, void* workspace,
    size_t workspace_bytes, void* stream);

/* Same name and arguments as the reference entry point
 * (warp/native/warp.h:563-602; ctypes argtypes warp/_src/context.py:7444-7466),
 * for callers that bind the reference library by name: topology only, runs
 * on the stream set with warp_amd_set_current_stream, allocates its scratch
 * stream-ordered, records bsr_nnz_event (a hipEvent_t) after the nnz copy.
 * Returns void like the reference; failures land in warp_amd_last_error(). */
void wp_bsr_matrix_from_triplets_device(
    int block_size, int scalar_size, int row_count, int col_count, int nnz, const int* tpl_nnz,
    const int* tpl_rows, const int* tpl_columns, const void* tpl_values, uint64_t scalar_zero_mask,
    bool masked_topology, int* tpl_block_offsets, int* tpl_block_indices, int* bsr_offsets,
    const int* bsr_row_counts, int* bsr_columns, int* bsr_nnz, void* bsr_nnz_event);

/* Sort and coalesce the active blocks of each row in place; replaces
 * wp_bsr_compress_inplace_device (warp/native/warp.h; host twin
 * warp/native/sparse.cpp:521-607, per-row semantics sparse_util.h:82-146) and
 * bsr_compress (warp/_src/sparse.py:1571-1672).  Active blocks: positions
 * [offsets[r], offsets[r] + row_counts[r]) (row_counts NULL: up to
 * offsets[r+1]) below T.  Negative columns are dropped, and with
 * prune_numerical_zeros blocks whose scalars all satisfy (bits & zero_mask)
 * == 0; duplicates are summed left to right in storage order (values NULL:
 * topology only; fp32 / fp64).  make_compact = 1: rows packed, offsets
 * rewritten, positions >= nnz get column -1 and zero values, *nnz_host = nnz
 * (pinned, asynchronous); 0: offsets kept, row_counts rewritten. */
size_t warp_amd_bsr_compress_workspace_size(int nrow, int T, int block_bytes);
int warp_amd_bsr_compress(int block_size, int scalar_size, int nrow, int ncol, int T, int prune_numerical_zeros,
                          uint64_t zero_mask, int make_compact, int* offsets, int* row_counts, int* columns,
                          void* values, int* nnz_host, void* workspace, size_t workspace_bytes, void* stream);

/* Triplets into the row capacity of a padded matrix (bsr_set_from_triplets
 * with topology="padded", warp/_src/sparse.py:1335-1420, whose scatter is
 * _bsr_set_from_triplets_padded_scatter_compact_topology, 1173-1208).
 * offsets[0..nrow] is read as capacity and left unchanged; row r gets its
 * sorted unique columns at offsets[r].. and their number in row_counts[r]; a
 * row that does not fit is left empty and raises *status (sticky) to 1
 * (BSR_STATUS_ROW_CAPACITY_EXCEEDED).  With values the first `capacity` blocks
 * are cleared and the sums written, duplicates added in ascending triplet
 * order; invalid and (zero_mask != 0) all-zero triplets are dropped. */
size_t warp_amd_bsr_topology_workspace_size(int nrow, int T, int block_bytes);
int warp_amd_bsr_set_from_triplets_padded(int block_size, int scalar_size, int nrow, int ncol, int T,
                                          const int* tpl_count, const int* rows, const int* cols,
                                          const void* tpl_values, uint64_t zero_mask, const int* offsets,
                                          int* row_counts, int* columns, void* values, int64_t capacity,
                                          int* status, void* workspace, size_t workspace_bytes, void* stream);

/* Triplets into an existing topology, left unchanged (bsr_set_from_triplets
 * with topology="masked", warp/_src/sparse.py:1422-1441 and the kernel
 * _bsr_set_from_triplets_masked_values, 1145-1170): the first `capacity`
 * blocks of values are cleared, then each active block (r, c) receives the sum
 * of the valid triplets at (r, c) in ascending triplet order (the reference
 * adds with atomics, in no fixed order).  No zero pruning. */
int warp_amd_bsr_set_from_triplets_masked(int block_size, int scalar_size, int nrow, int ncol, int T,
                                          const int* tpl_count, const int* rows, const int* cols,
                                          const void* tpl_values, const int* offsets, const int* row_counts,
                                          const int* columns, void* values, int64_t capacity, void* workspace,
                                          size_t workspace_bytes, void* stream);

/* Topology of the transpose of an nrow x ncol block matrix; replaces
 * wp_bsr_transpose_device / _host (warp/native/sparse.cu:1400-1492,
 * sparse.cpp:415-519; called from bsr_set_transpose,
 * warp/_src/sparse.py:2636-2833).  Active source blocks (positions < T) are
 * ordered by (column, row, position), none merged.  t_row_counts NULL:
 * compact, t_offsets[0..ncol] and t_columns / block_map[0..nnz) written.
 * Otherwise padded: t_offsets read as row capacity; t_row_counts, t_columns
 * and block_map written over the capacity (block_map -1 on unused slots); an
 * overflowing row is left empty and raises *status.  block_map[b] = source
 * position of transposed block b. */
size_t warp_amd_bsr_transpose_workspace_size(int ncol, int T);
int warp_amd_bsr_transpose(int nrow, int ncol, int T, const int* offsets, const int* row_counts, const int* columns,
                           int* t_offsets, int* t_row_counts, int* t_columns, int* block_map, int* status,
                           void* workspace, size_t workspace_bytes, void* stream);

/* Same name and arguments as the reference's wp_bsr_transpose_device
 * (warp/native/warp.h; bound in warp/_src/context.py), so its binding can
 * point here.  Stream from warp_amd_set_current_stream; scratch allocated
 * stream-ordered; errors through warp_amd_last_error(). */
void wp_bsr_transpose_device(int row_count, int col_count, int nnz, const int* bsr_offsets, const int* bsr_row_counts,
                             const int* bsr_columns, int* transposed_bsr_offsets, int* transposed_bsr_row_counts,
                             int* transposed_bsr_columns, int* src_block_indices, int* status);

/* dst[b] = scale * transpose(src[block_map[b]]) for b < min(count, *total)
 * with block_map[b] >= 0 (*total: device int, the transposed offsets[nrow]);
 * source blocks br x bc; fp32 (4) or fp64 (8).  Replaces the kernel
 * _bsr_transpose_values, warp/_src/sparse.py:2604-2633. */
int warp_amd_bsr_transpose_values(int scalar_size, int br, int bc, int64_t count, const int* total,
                                  const int* block_map, const void* src, double scale, void* dst, void* stream);

/* Masked transpose values (bsr_set_transpose with topology="masked"): each
 * active destination block (r, c) takes scale * transpose of source block
 * (c, r) when present (binary search in source row c).  Destination values
 * are not cleared here. */
int warp_amd_bsr_transpose_masked_values(int scalar_size, int br, int bc, int t_nrow, const int* t_offsets,
                                         const int* t_row_counts, const int* t_columns, int nrow, const int* offsets,
                                         const int* row_counts, const int* columns, const void* src, double scale,
                                         void* dst, void* stream);

/* Block coordinates of an assignment that reshapes blocks (replaces the
 * kernel _bsr_assign_list_blocks of bsr_assign, warp/_src/sparse.py:2258-2452):
 * each source block (positions < T) splits into src_subrows x src_subcols
 * pieces; destination blocks group dest_subrows x dest_subcols pieces.
 * Entry (p * src_subcols + subcol) * src_subrows + subrow receives the
 * destination (row, col) of that piece, (-1, -1) for inactive positions. */
int warp_amd_bsr_assign_list_blocks(int T, int src_subrows, int src_subcols, int dest_subrows, int dest_subcols,
                                    int nrow, const int* offsets, const int* row_counts, const int* columns,
                                    int* dest_rows, int* dest_cols, void* stream);

/* Copy every active source block's pieces into the existing destination
 * topology (replaces _bsr_assign_copy_blocks): value = (dest type)(scale *
 * source value); pieces without a destination block are skipped; destination
 * values are not cleared.  Scalar sizes 4 or 8. */
int warp_amd_bsr_assign_copy_blocks(int src_scalar_size, int dest_scalar_size, int T, int src_subrows,
                                    int src_subcols, int dest_subrows, int dest_subcols, int nrow, const int* offsets,
                                    const int* row_counts, const int* columns, int src_br, int src_bc,
                                    const void* src_values, double scale, const int* dest_offsets,
                                    const int* dest_row_counts, const int* dest_columns, int dest_br, int dest_bc,
                                    void* dest_values, void* stream);

/* Padded assignment with the same block shape (bsr_assign with
 * topology="padded", warp/_src/sparse.py:2331-2368, kernels
 * _bsr_assign_padded_row_ranges / _bsr_assign_padded_copy_blocks): each
 * source row's active blocks, in storage order, go to the start of the
 * destination row's capacity; a row that does not fit is left empty and
 * raises *status to 1.  src_values NULL: structure only. */
int warp_amd_bsr_assign_padded(int src_scalar_size, int dest_scalar_size, int nrow, int block_size,
                               const int* offsets, const int* row_counts, const int* columns, const void* src_values,
                               double scale, const int* dest_offsets, int* dest_row_counts, int* dest_columns,
                               void* dest_values, int* status, void* stream);

/* Padded assignment with a change of block shape, topology part (bsr_assign
 * with topology="padded", warp/_src/sparse.py:2201-2255 and 2346-2358,
 * kernel _bsr_assign_padded_reblock_topology, 2038-2083): per destination
 * row, the sorted destination columns that the covering source rows' blocks
 * overlap, written within the row's existing capacity; dest_row_counts set;
 * a row that does not fit is left empty and raises *status to 1.  Values
 * follow with warp_amd_bsr_assign_copy_blocks into zeroed storage. */
int warp_amd_bsr_assign_padded_reblock_topology(int src_subrows, int src_subcols, int dest_subrows, int dest_subcols,
                                                int src_nrow, int dest_nrow, int dest_ncol, const int* src_offsets,
                                                const int* src_row_counts, const int* src_columns,
                                                const int* dest_offsets, int* dest_row_counts, int* dest_columns,
                                                int* status, void* stream);

/* Row capacity of a reblocked padded copy as offsets[dest_nrow + 1] (bsr_copy
 * with topology="padded" and another block shape, warp/_src/sparse.py:
 * 2561-2588, kernel _bsr_copy_reblocked_capacity_counts, 2086-2108): each
 * destination row gets the capacity of its covering source rows times the
 * sub-columns per source block. */
size_t warp_amd_bsr_reblock_capacity_workspace_size(int dest_nrow);
int warp_amd_bsr_reblock_capacity(int src_subrows, int src_subcols, int dest_subrows, int dest_subcols, int src_nrow,
                                  int dest_nrow, const int* src_offsets, int* dest_offsets, void* workspace,
                                  size_t workspace_bytes, void* stream);

/* rows[p] = block row of storage position p < T, -1 where p holds no active
 * block (replaces the kernel of BsrMatrix.uncompress_rows,
 * warp/_src/sparse.py:164-193). */
int warp_amd_bsr_uncompress_rows(int nrow, const int* offsets, const int* row_counts, int T, int* rows, void* stream);

/* dst[i] = (dst type)(scale * src[i]) for i < n, scale in the source type:
 * the values of a same-shape compact assignment (replaces the block copy of
 * _bsr_assign_copy_blocks, warp/_src/sparse.py:2258-2452, when the
 * destination topology is the source's own). */
int warp_amd_copy_scaled_values(int src_scalar_size, int dst_scalar_size, int64_t n, const void* src, double scale,
                                void* dst, void* stream);

/* x[i] = 0 + beta * x[i] for i < n (fused multiply-add off): the values a
 * compact union-topology rebuild gives the blocks it keeps, so sums whose
 * topology is already present skip the rebuild bit for bit.  No reference
 * counterpart (the reference always rebuilds, sparse.py:3340-3617,
 * 4200-4640). */
int warp_amd_rescale_values(int scalar_size, int64_t n, void* x, double beta, void* stream);

/* *missing = number of pairs (rows[i], cols[i]), rows[i] >= 0, i < n, that
 * have no block in the matrix (offsets, row_counts, columns; sorted rows),
 * found by binary search as bsr_block_index does (sparse.py:1852-1900).
 * *missing is cleared by the call. */
int warp_amd_bsr_count_missing_blocks(int64_t n, const int* rows, const int* cols, const int* offsets,
                                      const int* row_counts, const int* columns, int* missing, void* stream);

/* dest[block (rows[i], cols[i])] += scale * src[i] for i < n; missing blocks
 * and rows[i] < 0 skipped (replaces _bsr_axpy_add_block,
 * warp/_src/sparse.py:3145-3163).  At most one source block per destination
 * block per call. */
int warp_amd_bsr_add_blocks(int scalar_size, int block_size, int64_t n, const int* rows, const int* cols,
                            const void* src, double scale, const int* dest_offsets, const int* dest_row_counts,
                            const int* dest_columns, void* dest_values, void* stream);

/* y := alpha x + beta y into y's row capacity (bsr_axpy with
 * topology="padded", warp/_src/sparse.py:3395-3478 and its kernels
 * _bsr_axpy_padded_count/_fill/_finalize, 3191-3311): sorted union per row,
 * merged in place from the back; a row that does not fit is left unchanged
 * with row_counts 0 and raises *status to 1. */
int warp_amd_bsr_axpy_padded(int scalar_size, int block_size, int nrow, double alpha, double beta,
                             const int* x_offsets, const int* x_row_counts, const int* x_columns,
                             const void* x_values, const int* y_offsets, int* y_row_counts, int* y_columns,
                             void* y_values, int* status, void* stream);

/* Symbolic sparse product x @ y (replaces bsr_mm_count_coeffs and its scan,
 * warp/_src/sparse.py:3620-3680, 4455-4475): mm_offsets[0..T] = exclusive
 * scan over the T x positions of the product blocks each lists, row_min as
 * the reference, *total_host = mm_offsets[T] (pinned, asynchronous,
 * optional).  deep_ranges = 1: a row with more products than the span of the
 * columns it reaches lists the whole span (the reference's compact rule). */
size_t warp_amd_bsr_mm_count_workspace_size(int T);
int warp_amd_bsr_mm_count(int T, int x_nrow, int y_nrow, int y_ncol, int deep_ranges, const int* x_offsets,
                          const int* x_row_counts, const int* x_columns, const int* y_offsets,
                          const int* y_row_counts, const int* y_columns, int* row_min, int* mm_offsets,
                          int* total_host, void* workspace, size_t workspace_bytes, void* stream);

/* Product triplets (replaces _bsr_mm_list_coeffs, warp/_src/sparse.py:3683-3735):
 * rows/cols[base + i], i < n, = (row, column) of product block i; (-1, -1)
 * beyond mm_offsets[T]. */
int warp_amd_bsr_mm_list(int n, int base, int T, int x_nrow, const int* x_offsets, const int* x_row_counts,
                         const int* x_columns, const int* y_offsets, const int* y_columns, const int* row_min,
                         const int* mm_offsets, int* rows, int* cols, void* stream);

/* Product values into an existing topology (replaces _bsr_mm_compute_values,
 * warp/_src/sparse.py:3764-3800): z(r, c) = (0 + beta * z(r, c)) + alpha * sum
 * over x row r's active blocks k, in storage order, of x_k @ y(x_columns[k],
 * c).  x blocks br x bk, y bk x bc (br * bc <= 64); products and sums rounded
 * one by one (no fused multiply-add). */
int warp_amd_bsr_mm_values(int scalar_size, int br, int bk, int bc, double alpha, double beta, int T, int z_nrow,
                           const int* z_offsets, const int* z_row_counts, const int* z_columns, void* z_values,
                           int y_nrow, const int* x_offsets, const int* x_row_counts, const int* x_columns,
                           const void* x_values, const int* y_offsets, const int* y_row_counts, const int* y_columns,
                           const void* y_values, void* stream);

/* Tiled product values (bsr_mm with tile_size > 0; replaces the tiled-outer
 * values kernel, warp/_src/sparse.py:3813-3909, and its dispatch, 4101-4144):
 * `tile` lanes per z block (a power of two in [1, 64]); the x row is read as
 * the flat list of (x block k, inner index m) pairs, m fastest; lane g adds
 * the rank-one updates x_k[:, m] y(x_columns[k], c)[m, :] of pairs g, g+tile,
 * ... (each product rounded, then added); the lanes combine by an xor
 * butterfly; z(r, c) = (0 + beta * z(r, c)) + alpha * sum. */
int warp_amd_bsr_mm_values_tiled(int scalar_size, int br, int bk, int bc, double alpha, double beta, int T, int tile,
                                 int z_nrow, const int* z_offsets, const int* z_row_counts, const int* z_columns,
                                 void* z_values, int y_nrow, const int* x_offsets, const int* x_row_counts,
                                 const int* x_columns, const void* x_values, const int* y_offsets,
                                 const int* y_row_counts, const int* y_columns, const void* y_values, void* stream);

/* Block-diagonal topology: offsets[r] = min(r, ndiag), columns[r] = r for
 * r < ndiag (replaces _bsr_set_diag_kernel of bsr_set_diag,
 * warp/_src/sparse.py:2939-2995). */
int warp_amd_bsr_diag_topology(int nrow, int ndiag, int* offsets, int* columns, void* stream);

/* y = alpha x + beta y over n scalars (replaces fem array_axpy,
 * warp/_src/fem/linalg.py:425-461). */
int warp_amd_axpby(int scalar_size, int64_t n, double alpha, const void* x, double beta, void* y, void* stream);

/* Dirichlet projector normalization (replaces normalize_dirichlet_projector's
 * kernels, warp/_src/fem/dirichlet.py:12-66, 127-199): each diagonal block
 * (block_dim <= 8) becomes sum v v^T over its eigenpairs with eigenvalue
 * above 1e-6 trace, the row's fixed value (NULL: none) sum (v.b)/lambda v.
 * FP64 Jacobi eigen decomposition. */
int warp_amd_normalize_dirichlet_projector(int scalar_size, int block_dim, int nrow, const int* offsets,
                                           const int* row_counts, const int* columns, void* values,
                                           void* fixed_values, void* stream);

/* x[0..n) *= alpha, fp32 (4) or fp64 (8) (bsr_scale's value kernel). */
int warp_amd_scale_values(int scalar_size, int64_t n, void* x, double alpha, void* stream);

/* ------------------------------------------------------------------------ */
/* SpMV, diagonal, Jacobi                                                   */
/* ------------------------------------------------------------------------ */

/* y = alpha * A x + beta * y for a BSR matrix with block_rows x block_cols
 * blocks (row_counts NULL = compact).  Replaces the JIT kernels bsr_mv_kernel
 * and bsr_mv_tiled_kernel (warp/_src/sparse.py:4627-4715, launched from
 * bsr_mv, 4792-4913).  alpha == 0 leaves x unread, beta == 0 leaves y unread.
 * nnz_hint: stored-block count (or upper bound) used to pick lanes per row;
 * <= 0 for a default.  1x1 and 3x3 blocks use the vector-BSR kernels, other
 * shapes a one-thread-per-scalar-row kernel. */
int warp_amd_bsr_mv_f32(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const float* values, int64_t nnz_hint, const float* x, float* y, float alpha, float beta,
    void* stream);
int warp_amd_bsr_mv_f64(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const double* values, int64_t nnz_hint, const double* x, double* y, double alpha,
    double beta, void* stream);

/* y = alpha * A^T x + beta * y (replaces bsr_mv with transpose=True and its
 * bsr_mv_transpose_kernel, warp/_src/sparse.py:4718-4747, 4880-4895).  Shapes
 * are A's: x has nrow * block_rows entries, y has ncol * block_cols.  Like
 * the reference, an atomic scatter: the last bits of y vary between runs. */
int warp_amd_bsr_mv_transpose_f32(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const float* values, int64_t nnz_hint, const float* x, float* y, float alpha, float beta,
    void* stream);
int warp_amd_bsr_mv_transpose_f64(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const double* values, int64_t nnz_hint, const double* x, double* y, double alpha,
    double beta, void* stream);

/* y = alpha * A^T x + beta * y without atomics, deterministic (the variant
 * of the reference's atomic bsr_mv_transpose_kernel, warp/_src/sparse.py:
 * 4718-4747, for structurally symmetric patterns): A compact, square, blocks
 * block_size x block_size (1 or 3), every block (r, c) with a mirror (c, r).
 * Output row c sums, over its own blocks (c, r), the mirror block (r, c)^T
 * x[r] (found by binary search); per warp_amd_bsr_mv_lanes lanes, an xor
 * butterfly.  Blocks without a mirror contribute nothing: check the pattern
 * first (warp_amd_bsr_count_missing_blocks of the transposed pairs). */
int warp_amd_bsr_mv_transpose_gather_f32(int nrow, int block_size, const int* offsets, const int* columns,
                                         const float* values, int64_t nnz_hint, const float* x, float* y,
                                         float alpha, float beta, void* stream);
int warp_amd_bsr_mv_transpose_gather_f64(int nrow, int block_size, const int* offsets, const int* columns,
                                         const double* values, int64_t nnz_hint, const double* x, double* y,
                                         double alpha, double beta, void* stream);

/* Same with explicit kernel shape: `lanes` per block row (power of two in
 * [1, 64]; the reference's bsr_mv `tile_size` plays this role) and
 * `rows_per_group` consecutive rows per lane group (1, 2 or 4); <= 0 picks the
 * default.  Results depend on `lanes` only. */
int warp_amd_bsr_mv_ex_f32(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const float* values, int64_t nnz_hint, const float* x, float* y, float alpha, float beta,
    int lanes, int rows_per_group, void* stream);
int warp_amd_bsr_mv_ex_f64(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const double* values, int64_t nnz_hint, const double* x, double* y, double alpha,
    double beta, int lanes, int rows_per_group, void* stream);

/* The same product with the matrix's arguments packed once into a host
 * struct (no reference counterpart: a binding that calls bsr_mv on the same
 * matrix many times fills it once and passes 6 arguments per call instead of
 * 16, which is most of the host cost of a small product).  scalar_size 4:
 * float values / vectors, 8: double; alpha and beta are rounded to the
 * scalar type.  The struct is read during the call only.  sweep bit 0
 * reverses the order in which the workgroups take the block rows (1x1 and
 * 3x3 blocks; the result is bit-identical either way): a caller that flips
 * it between consecutive products with one matrix reads first, in each
 * product, the rows the previous one read last, while they are still in
 * the GPU's memory-side cache. */
typedef struct warp_amd_bsr_mv_args {
    int nrow, ncol, block_rows, block_cols;
    const int* offsets;    /* device */
    const int* row_counts; /* device, NULL = compact */
    const int* columns;    /* device */
    const void* values;    /* device */
    int64_t nnz_hint;
    int lanes, rows_per_group, scalar_size, sweep;
} warp_amd_bsr_mv_args;
int warp_amd_bsr_mv_packed(const warp_amd_bsr_mv_args* args, const void* x, void* y, double alpha, double beta,
                           void* stream);

/* Lanes per block row G the vector-BSR kernels (bsr_mv and the fused CG
 * kernels) use for this shape and nnz_hint; 1 for the generic kernel.  Each
 * lane sums blocks g, g+G, ... with fused multiply-adds, then the group
 * combines by an xor butterfly: callers can reproduce the bits. */
int warp_amd_bsr_mv_lanes(int nrow, int block_rows, int block_cols, int64_t nnz_hint);

/* out[r] = scale * A[r, r] (zero block where absent), r < min(nrow, ncol).
 * Replaces _bsr_get_diag_kernel (warp/_src/sparse.py:2871-2924). */
int warp_amd_bsr_get_diag_f32(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const float* values, float scale, float* out, void* stream);
int warp_amd_bsr_get_diag_f64(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const double* values, double scale, double* out, void* stream);

/* Jacobi data straight from the matrix.  mode 0: out[nrow * br] = 1/d of the
 * diagonal entries of each diagonal block (1/|d| with use_abs, 1 where d == 0)
 * -- replaces bsr_get_diag + _extract_inverse_diagonal_{blocked,scalar}
 * (warp/_src/optim/linear.py:256-280, 2290-2320).  mode 1 (extension, no
 * reference counterpart): out[nrow * br * br] = inverse of each diagonal
 * block (block Jacobi; singular blocks fall back to mode 0 values). */
int warp_amd_bsr_jacobi_f32(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const float* values, int use_abs, int mode, float* out, void* stream);
int warp_amd_bsr_jacobi_f64(
    int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,
    const int* columns, const double* values, int use_abs, int mode, double* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Conjugate gradient                                                       */
/* ------------------------------------------------------------------------ */

typedef int (*warp_amd_cg_callback)(int iteration, double residual_norm, double abs_tolerance, void* user);

size_t warp_amd_cg_workspace_size(int64_t n, int scalar_size);
/* byte offset inside the CG workspace of: 0 the iteration count (int32),
 * 1 r.r (double), 2 atol^2 (double) -- the device results of check_every == 0 */
size_t warp_amd_cg_state_offset(int which);

/* Preconditioned CG on a compact BSR matrix with 1x1 or 3x3 blocks.
 * Replaces CG._run / cg (warp/_src/optim/linear.py:893-1121) with its kernels
 * _cg_kernel_1/2 (2332-2373), TiledDot (309-612), _initialize_tolerance
 * (2051-2083) and the host loop _run_solver_loop (2182-2243).
 *
 * prec_kind 0 none, 1 diagonal (prec = n scalars), 2 block (prec = nrow
 * block_size^2 scalars).  tol / atol: NaN = None (defaults of
 * linear.py:2010-2032).  maxiter <= 0 = n.  check_every > 0: residual checks
 * and callback every check_every iterations, results in *iterations,
 * *residual_norm, *abs_tolerance (host).  check_every == 0: no host checks;
 * the convergence flag stays on the device and results are read from the
 * workspace (warp_amd_cg_state_offset).  restart > 0: recompute r = b - A x
 * every `restart` iterations.  use_graph (use_cuda_graph of
 * linear.py:1035-1049): for systems of at most 1 MiB per vector, whose
 * iterations are launch-bound, run on a private stream and replay the cycles
 * between two residual checks as one HIP graph; larger systems and callers
 * whose stream is being captured launch directly (same iterates either
 * way).  x: initial guess in, solution out. */
int warp_amd_cg_solve_f32(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    int prec_kind, const float* prec, const float* b, float* x, double tol, double atol, int maxiter,
    int check_every, int restart, int use_graph, warp_amd_cg_callback callback, void* callback_user,
    void* workspace, size_t workspace_bytes, int* iterations, double* residual_norm, double* abs_tolerance,
    void* stream);
int warp_amd_cg_solve_f64(
    int nrow, int block_size, const int* offsets, const int* columns, const double* values, int64_t nnz_hint,
    int prec_kind, const double* prec, const double* b, double* x, double tol, double atol, int maxiter,
    int check_every, int restart, int use_graph, warp_amd_cg_callback callback, void* callback_user,
    void* workspace, size_t workspace_bytes, int* iterations, double* residual_norm, double* abs_tolerance,
    void* stream);

/* Generic CG building blocks for caller-supplied operators.
 * *out = sum a[i] b[i] accumulated in FP64, deterministic (replaces TiledDot
 * block_dot_kernel / block_sum_kernel, linear.py:580-612); the workspace
 * (warp_amd_dot_workspace_size) must be zero-filled before its first use. */
size_t warp_amd_dot_workspace_size(int64_t n);
int warp_amd_dot_f32(int64_t n, const float* a, const float* b, double* out, void* workspace, size_t workspace_bytes,
                     void* stream);
int warp_amd_dot_f64(int64_t n, const double* a, const double* b, double* out, void* workspace,
                     size_t workspace_bytes, void* stream);
/* *atol_sq = max(rtol * sqrt(*bb), atol)^2 with the tolerance defaults
 * (linear.py:2010-2060); device scalars. */
int warp_amd_cg_tolerance(const double* bb, double tol, double atol, int is_f64, double* atol_sq, void* stream);
/* alpha = rr > atol_sq ? rz / pAp : 0; x += alpha p; r -= alpha Ap (_cg_kernel_1, linear.py:2332-2353) */
int warp_amd_cg_step_xr_f32(int64_t n, float* x, float* r, const float* p, const float* Ap, const double* rr,
                            const double* atol_sq, const double* rz, const double* pAp, void* stream);
int warp_amd_cg_step_xr_f64(int64_t n, double* x, double* r, const double* p, const double* Ap, const double* rr,
                            const double* atol_sq, const double* rz, const double* pAp, void* stream);
/* beta = rr > atol_sq ? rz_new / rz_old : 0; p = z + beta p (_cg_kernel_2, linear.py:2355-2373) */
int warp_amd_cg_step_p_f32(int64_t n, const float* z, float* p, const double* rr, const double* atol_sq,
                           const double* rz_new, const double* rz_old, void* stream);
int warp_amd_cg_step_p_f64(int64_t n, const double* z, double* p, const double* rr, const double* atol_sq,
                           const double* rz_new, const double* rz_old, void* stream);
/* Batched solves (LinearOperator.batch_offsets, linear.py:33-131): scalar dofs
 * [batch_offsets[b], batch_offsets[b + 1]) form subproblem b < nb; dofs
 * outside [batch_offsets[0], batch_offsets[nb]) are inactive.
 * warp_amd_seg_dot_plan computes chunk_start[0..nb] once per layout
 * (workspace: warp_amd_seg_dot_plan_workspace_size); warp_amd_seg_dot_* then
 * gives out[b] = sum over subproblem b of a[i] b[i] in FP64, each subproblem
 * cut into chunks of 1024 dofs summed in a fixed order (replaces the segmented
 * TiledDot, linear.py:309-577); partials holds warp_amd_seg_dot_max_chunks(n,
 * nb) doubles, max_chunks being that value. */
int warp_amd_seg_dot_max_chunks(int64_t n, int nb);
size_t warp_amd_seg_dot_plan_workspace_size(int nb);
int warp_amd_seg_dot_plan(int nb, const int* batch_offsets, int* chunk_start, void* workspace, size_t workspace_bytes,
                          void* stream);
int warp_amd_seg_dot_f32(int64_t n, int nb, const int* batch_offsets, const int* chunk_start, int max_chunks,
                         const float* a, const float* b, double* partials, double* out, void* stream);
int warp_amd_seg_dot_f64(int64_t n, int nb, const int* batch_offsets, const int* chunk_start, int max_chunks,
                         const double* a, const double* b, double* partials, double* out, void* stream);
/* atol_sq[b] = max(rtol * sqrt(bb[b]), atol)^2 per subproblem
 * (_initialize_tolerance, linear.py:2051-2083). */
int warp_amd_cg_tolerance_batched(int nb, const double* bb, double tol, double atol, int is_f64, double* atol_sq,
                                  void* stream);
/* _cg_kernel_1 / _cg_kernel_2 with batches (linear.py:2332-2373): entry e of
 * dpe scalars belongs to the subproblem of its first dof (_find_entry_batch,
 * linear.py:2034-2048) and uses that subproblem's alpha / beta; inactive
 * entries are left unchanged. */
int warp_amd_cg_step_xr_batched_f32(int64_t nent, int dpe, int nb, const int* batch_offsets, float* x, float* r,
                                    const float* p, const float* Ap, const double* rr, const double* atol_sq,
                                    const double* rz, const double* pAp, void* stream);
int warp_amd_cg_step_xr_batched_f64(int64_t nent, int dpe, int nb, const int* batch_offsets, double* x, double* r,
                                    const double* p, const double* Ap, const double* rr, const double* atol_sq,
                                    const double* rz, const double* pAp, void* stream);
int warp_amd_cg_step_p_batched_f32(int64_t nent, int dpe, int nb, const int* batch_offsets, const float* z, float* p,
                                   const double* rr, const double* atol_sq, const double* rz_new,
                                   const double* rz_old, void* stream);
int warp_amd_cg_step_p_batched_f64(int64_t nent, int dpe, int nb, const int* batch_offsets, const double* z,
                                   double* p, const double* rr, const double* atol_sq, const double* rz_new,
                                   const double* rz_old, void* stream);
/* z = alpha (d o x) + beta y (_diag_mv_kernel, linear.py:2271-2287) */
int warp_amd_diag_mv_f32(int64_t n, const float* d, const float* x, const float* y, float* z, float alpha, float beta,
                         void* stream);
int warp_amd_diag_mv_f64(int64_t n, const double* d, const double* x, const double* y, double* z, double alpha,
                         double beta, void* stream);
/* z = alpha A x + beta y, A dense row-major m x k (_dense_mv_kernel, linear.py:2246-2268) */
int warp_amd_dense_mv_f32(int m, int k, const float* A, const float* x, const float* y, float* z, float alpha,
                          float beta, void* stream);
int warp_amd_dense_mv_f64(int m, int k, const double* A, const double* x, const double* y, double* z, double alpha,
                          double beta, void* stream);

/* ------------------------------------------------------------------------ */
/* Mixed precision: FP64 iterative refinement around the FP32 CG            */
/* (no reference counterpart: the reference solves in one precision)        */
/* ------------------------------------------------------------------------ */

size_t warp_amd_bsr_residual_workspace_size(int nrow);
/* r64 = b - A x, FP64 accumulation of the FP32 matrix values (1x1 or 3x3
 * blocks, compact), FP64 x and b; r32 = (float) r64; *rr (device) = r64.r64
 * summed in a fixed order. */
int warp_amd_bsr_residual_f32_f64(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    const double* x, const double* b, double* r64, float* r32, double* rr, void* workspace, size_t workspace_bytes,
    void* stream);
/* warp_amd_bsr_residual_f32_f64 without a workspace.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The partials
 * live in a device array of the library (65,536 doubles, 512 KB, held by
 * every GPU the library runs on until the process ends): nrow is limited to
 * 65,536 residual workgroups (16,777,216 scalar rows, 1,048,576 rows of 3x3
 * blocks at the default lanes per row; larger sizes return an error before
 * any launch).  Calls that overlap on one GPU race on that array.  The first
 * call on a device looks up the array's address on that device. */
int warp_amd_bsr_residual_f32_f64_noalloc(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    const double* x, const double* b, double* r64, float* r32, double* rr, void* stream);
/* x64 += alpha * (double) d32 */
int warp_amd_axpy_f32_to_f64(int64_t n, double alpha, const float* d, double* x, void* stream);
/* y64 = A x64 (FP32 values promoted, FP64 accumulation); *xy (device) = x.y.
 * Workspace: warp_amd_bsr_residual_workspace_size(nrow). */
int warp_amd_bsr_mv_f32_f64(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    const double* x, double* y, double* xy, void* workspace, size_t workspace_bytes, void* stream);

/* Flexible CG pieces (Notay, SIAM J. Sci. Comput. 22(4), 2000; truncation 1),
 * FP64 vectors, every scalar on the device so an iteration can be captured in
 * a HIP graph.  scale = sqrt(*ss / n), inv_scale = 1 / scale. */
int warp_amd_rms_scales(const double* ss, int64_t n, double* scale, double* inv_scale, void* stream);
/* p = z - (*zq / *pq_prev) p_prev */
int warp_amd_fcg_direction(int64_t n, const double* z, const double* p_prev, const double* zq, const double* pq_prev,
                           double* p, void* stream);
/* alpha = *pr / *pq; x += alpha p; r -= alpha q; *rr = r.r (device, fixed-order
 * sum); with w non-null also w = (float)(omega * dinv * r) */
size_t warp_amd_fcg_update_workspace_size(int64_t n);
int warp_amd_fcg_update(int64_t n, const double* pr, const double* pq, const double* p, const double* q, double* x,
                        double* r, double* rr, const double* dinv, double omega, float* w, void* workspace,
                        size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Learned preconditioner: 3D U-Net building blocks (no reference           */
/* counterpart).  Fields on an n^3 grid are channels-last FP32 arrays       */
/* [voxel][channel], voxel = x + n (y + n z); convolutions 3x3x3, stride 1, */
/* zero padding, weights W[tap][ci][co], tap = (dz+1)*9 + (dy+1)*3 + dx+1.  */
/* ------------------------------------------------------------------------ */

/* y (+)= conv3(x) + b, ReLU when relu != 0 */
int warp_amd_conv3_fwd_f32(int n, int cin, int cout, const float* x, const float* W, const float* b, float* y,
                           int relu, int accumulate, void* stream);
/* dx (+)= conv3^T(dy) */
int warp_amd_conv3_bwd_data_f32(int n, int cin, int cout, const float* dy, const float* W, float* dx,
                                int accumulate, void* stream);
size_t warp_amd_conv3_bwd_weight_workspace_size(int n, int cin, int cout);
/* dW = sum_v x[v + d] dy[v], db = sum_v dy[v]; chunked, fixed-order sums */
int warp_amd_conv3_bwd_weight_f32(int n, int cin, int cout, const float* x, const float* dy, float* dW, float* db,
                                  void* workspace, size_t workspace_bytes, void* stream);
/* out = dy * (y > 0) */
int warp_amd_relu_mask_f32(int64_t count, const float* y, const float* dy, float* out, void* stream);
/* 2x2x2 average pooling (n even) and its adjoint */
int warp_amd_pool2_f32(int n, int ch, const float* x, float* y, void* stream);
int warp_amd_pool2_bwd_f32(int n, int ch, const float* dy, float* dx, int accumulate, void* stream);
/* nearest 2x upsampling of an n^3 field plus an optional fine skip field, and its adjoint */
int warp_amd_up2_add_f32(int n, int ch, const float* x, const float* skip, float* y, void* stream);
int warp_amd_up2_bwd_f32(int n, int ch, const float* dy, float* dx, void* stream);
/* dout = (out - target) / count; *loss = 0.5 mean((out - target)^2) (device) */
size_t warp_amd_mse_workspace_size(int64_t count);
int warp_amd_mse_loss_f32(int64_t count, const float* out, const float* target, float* dout, double* loss,
                          void* workspace, size_t workspace_bytes, void* stream);
/* Adam step; replaces adam_step_kernel_float (warp/_src/optim/adam.py:30-47) */
int warp_amd_adam_step_f32(int64_t count, const float* g, float* m, float* v, float* param, float lr, float beta1,
                           float beta2, float t, float eps, void* stream);
/* y = alpha * x + beta * y (FP32; gradient accumulation over a batch) */
int warp_amd_axpby_f32(int64_t count, float alpha, const float* x, float beta, float* y, void* stream);
/* network input (r * inv_scale, standardized log k) and output z = scale * out; scales on the device */
int warp_amd_unet_pack_input(int64_t nvox, const double* r, const float* logk_std, const double* inv_scale,
                             float* inp, void* stream);
int warp_amd_unet_unpack_output(int64_t nvox, const float* out, const double* scale, double* z, void* stream);

/* Smoothing around the learned correction (no reference counterpart): one
 * damped Jacobi sweep z_out = z + omega * dinv * (r - A z) for a 1x1 compact
 * BSR matrix with FP32 values and FP64 vectors (z_out must not alias z), and
 * z += mult * (*scale) * out for the network output. */
int warp_amd_jacobi_sweep_f32_f64(int nrow, const int* offsets, const int* columns, const float* values,
                                  int64_t nnz_hint, const double* z, const double* r, const double* dinv, double omega,
                                  double* z_out, void* stream);
int warp_amd_unet_unpack_add(int64_t n, const float* out, const double* scale, double mult, double* z, void* stream);
/* Fused flexible-CG steps for a scalar (1x1) compact BSR matrix with FP32
 * values and FP64 vectors, one damped Jacobi sweep before and after the
 * network correction (no reference counterpart).  w = (float)(omega D^-1 r)
 * is the first sweep, kept current by warp_amd_fcg_update; the
 * preconditioned z is FP32.  Scalars are device pointers; workspace >=
 * warp_amd_fcg_fused_workspace_size(nrow).
 *   weight:     w = (float)(omega D^-1 r)
 *   presmooth:  r1 = r - A w; *rr1 = r1.r1, *scale = rms(r1), *inv_scale = 1 / rms(r1)
 *   postsmooth: z2 = w + gain * (*scale) * out; z = (float)(z2 + omega D^-1 (r - A z2)); *zq = z.q_prev
 *   direction_spmv: p = z - (*zq / *pq_prev) p_prev; q = A p; *pq = p.q; *pr = p.r */
size_t warp_amd_fcg_fused_workspace_size(int nrow);
int warp_amd_fcg_weight(int64_t n, const double* dinv, const double* r, double omega, float* w, void* stream);
int warp_amd_fcg_presmooth(int nrow, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
                           const double* r, const float* w, double* r1, double* rr1, double* scale, double* inv_scale,
                           void* workspace, size_t workspace_bytes, void* stream);
int warp_amd_fcg_postsmooth(int nrow, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
                            const double* r, const double* dinv, double omega, const float* w, const float* out,
                            const double* scale, double gain, const double* q_prev, float* z, double* zq,
                            void* workspace, size_t workspace_bytes, void* stream);
int warp_amd_fcg_direction_spmv(int nrow, const int* offsets, const int* columns, const float* values,
                                int64_t nnz_hint, const float* z, const double* p_prev, const double* zq,
                                const double* pq_prev, const double* r, double* p, double* q, double* pq, double* pr,
                                void* workspace, size_t workspace_bytes, void* stream);
/* Training loss of the learned preconditioner (no reference counterpart): the
 * one-step energy reduction L = 1 - (r.z)^2 / ((r.e)(z.Az)) of a direction z
 * for A e = r.  sc (device) = {r.z, z.Az, r.e}; writes g = dL/dz and *loss. */
int warp_amd_energy_loss_grad(int64_t n, const double* r, const double* az, const double* sc, double* g, double* loss,
                              void* stream);
/* out32 = mult * (*scale) * x64 */
int warp_amd_scale_to_f32(int64_t n, const double* x, const double* scale, double mult, float* out, void* stream);

/* BF16 MFMA 3x3x3 convolution (no reference counterpart).  Activations are
 * channels-last BF16 [m^3][c]; weights are FP32 masters W[27][cin][cout]
 * swizzled once into BF16 MFMA operand tiles (transpose_flip = 1 builds the
 * backward-data operator of a (cout -> cin) layer).  in_mode 1 adds the 2x
 * nearest upsampling of `coarse` to x; flags: 1 ReLU, 2 also write the 2x2x2
 * average pool, 4 FP32 output yf, 8 accumulate into yf.  Channel pairs:
 * 2-16, 16-32, 32-64, 64-64, 64-32, 32-16, 16-1, 16-16, 1-16, 16-2. */
size_t warp_amd_conv3_mfma_weight_size(int cin, int cout);
int warp_amd_conv3_mfma_swizzle(int cin, int cout, int transpose_flip, const float* W, void* wsw, void* stream);
int warp_amd_conv3_mfma(int m, int cin, int cout, int in_mode, const void* x, const void* coarse, const void* wsw,
                        const float* bias, int flags, void* y, float* yf, void* pooled, void* stream);
/* FP32 <-> BF16 (round to nearest even) */
int warp_amd_f32_to_bf16(int64_t count, const float* in, void* out, void* stream);
int warp_amd_bf16_to_f32(int64_t count, const void* in, float* out, void* stream);
/* BF16 network input [nvox][2] = (r * inv_scale, standardized log k) */
int warp_amd_unet_pack_input_bf16(int64_t nvox, const double* r, const float* logk_std, const double* inv_scale,
                                  void* out, void* stream);
/* Weight and bias gradients of a BF16 (cin -> cout) convolution on MFMA (no
 * reference counterpart; training of the learned preconditioner):
 * dW[27][cin][cout] = sum_v x~[v + d_t][ci] dy[v][co], db[cout] = sum_v dy[v][co],
 * both overwritten.  x: the layer's BF16 input (in_mode 1: x + upsample(coarse),
 * rounded as the forward rounds it); dy: FP32 [m^3][cout], rounded to BF16; FP32
 * accumulation; per-workgroup partials in the workspace added in a fixed order.
 * Channel pairs: 2-16, 16-32, 32-64, 64-64, 64-32, 32-16, 16-1. */
size_t warp_amd_conv3_mfma_wgrad_workspace_size(int m, int cin, int cout);
int warp_amd_conv3_mfma_wgrad(int m, int cin, int cout, int in_mode, const void* x, const void* coarse,
                              const float* dy, float* dW, float* db, void* workspace, size_t workspace_bytes,
                              void* stream);
/* out = y > 0 ? dy : 0 for a BF16 y (ReLU backward of a BF16 forward); dy, out FP32 */
int warp_amd_relu_mask_bf16(int64_t count, const void* y, const float* dy, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WARP_AMD_H */
