// Whole-matrix BSR operations around the solver: assignment with block
// reshaping and scalar conversion (bsr_assign / bsr_copy) and value scaling.
//
// Replaces the JIT kernels _bsr_assign_list_blocks and _bsr_assign_copy_blocks
// (warp/_src/sparse.py, launched from bsr_assign, 2258-2452) and
// _bsr_scale_kernel (bsr_scale).  The compact topology of an assignment is
// built by the triplet assembly (one triplet per destination sub-block).
#include "common.h"

#include <climits>

namespace wa {
namespace {

constexpr int kBlock = 256;

// Row of storage position p, -1 when p holds no active block (reference
// bsr_row_index with row counts, warp/_src/sparse.py:1794-1828).
__device__ __forceinline__ int row_of_position(const int* offsets, const int* row_counts, int nrow, int p)
{
    if (nrow <= 0 || p >= offsets[nrow])
        return -1;
    int lo = 0, hi = nrow - 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (offsets[mid + 1] <= p)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (row_counts && p >= offsets[lo] + row_counts[lo])
        return -1;
    return lo;
}

// Position of block (row, col) in a row-sorted topology, -1 if absent
// (reference bsr_block_index, warp/_src/sparse.py:1852-1900).
__device__ __forceinline__ int find_block(const int* offsets, const int* row_counts, const int* columns, int row,
                                          int col)
{
    if (row < 0)
        return -1;
    const int beg = offsets[row];
    const int end = row_counts ? beg + row_counts[row] : offsets[row + 1];
    int lo = beg, hi = end;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (columns[mid] < col)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < end && columns[lo] == col ? lo : -1;
}

// Active end of row r, clipped to its capacity.
__device__ __forceinline__ int row_end(const int* offsets, const int* row_counts, int r)
{
    const int cap_end = offsets[r + 1];
    if (!row_counts)
        return cap_end;
    const int e = offsets[r] + row_counts[r];
    return e < cap_end ? e : cap_end;
}

// Row-parallel launches: one wavefront per row, kRowsPerBlock rows per
// workgroup; a row's storage positions are walked directly (no search for the
// row of a position).
constexpr int kRowsPerBlock = 4;

__device__ __forceinline__ int wave_row() { return blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave); }

struct Reblock {
    int src_subrows, src_subcols, dest_subrows, dest_subcols;
};

// One wavefront per source row, a lane per (position, sub-row, sub-column):
// the destination block coordinates of that sub-block at entry
// (p * src_subcols + subcol) * src_subrows + subrow.  Entries of inactive
// positions were set to -1 beforehand.
__global__ __launch_bounds__(kWave * kRowsPerBlock) void assign_list_blocks_kernel(int T, Reblock rb, int nrow,
                                                                                  const int* __restrict__ offsets,
                                                                                  const int* __restrict__ row_counts,
                                                                                  const int* __restrict__ columns,
                                                                                  int* __restrict__ dest_rows,
                                                                                  int* __restrict__ dest_cols)
{
    const int row = wave_row();
    if (row >= nrow)
        return;
    const int per = rb.src_subrows * rb.src_subcols;
    const int beg = offsets[row];
    const int end = min(row_end(offsets, row_counts, row), T);
    for (int t = (int)(threadIdx.x % kWave); t < (end - beg) * per; t += kWave) {
        const int p = beg + t / per;
        const int rem = t % per;
        const int subcol = rem % rb.src_subcols;
        const int subrow = rem / rb.src_subcols;
        const int col = columns[p];
        if (col < 0)
            continue;
        const int64_t o = ((int64_t)p * rb.src_subcols + subcol) * rb.src_subrows + subrow;
        dest_rows[o] = (row * rb.src_subrows + subrow) / rb.dest_subrows;
        dest_cols[o] = (col * rb.src_subcols + subcol) / rb.dest_subcols;
    }
}

// One wavefront per source row, a lane per (position, sub-row, sub-column):
// copies the sub-block into its destination block, converted to the
// destination scalar type after the scale is applied in the source type.
template <typename SS, typename DS>
__global__ __launch_bounds__(kWave * kRowsPerBlock) void assign_copy_blocks_kernel(
    int T, Reblock rb, int nrow, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, int src_br, int src_bc, const SS* __restrict__ src, SS scale,
    const int* __restrict__ d_offsets, const int* __restrict__ d_row_counts, const int* __restrict__ d_columns,
    int dst_br, int dst_bc, DS* __restrict__ dst)
{
    const int row = wave_row();
    if (row >= nrow)
        return;
    const int per = rb.src_subrows * rb.src_subcols;
    const int beg = offsets[row];
    const int end = min(row_end(offsets, row_counts, row), T);
    if (per == 1 && rb.dest_subrows == 1 && rb.dest_subcols == 1) {
        // same block shape: lanes over (block, scalar), scalars fastest, so
        // loads and stores are contiguous; the destination block usually sits
        // at the same rank in its row (copies of a sorted topology)
        const int bs = src_br * src_bc;
        const int dbeg = d_offsets[row];
        const int dend = row_end(d_offsets, d_row_counts, row);
        for (int t = (int)(threadIdx.x % kWave); t < (end - beg) * bs; t += kWave) {
            const int k = t / bs;
            const int e = t - k * bs;
            const int p = beg + k;
            const int col = columns[p];
            if (col < 0)
                continue;
            int b = dbeg + k;
            if (!(b < dend && d_columns[b] == col))
                b = find_block(d_offsets, d_row_counts, d_columns, row, col);
            if (b >= 0)
                dst[(size_t)b * bs + e] = (DS)(scale * src[(size_t)p * bs + e]);
        }
        return;
    }
    for (int t = (int)(threadIdx.x % kWave); t < (end - beg) * per; t += kWave) {
        const int p = beg + t / per;
        const int rem = t % per;
        const int subcol = rem % rb.src_subcols;
        const int subrow = rem / rb.src_subcols;
        const int col = columns[p];
        if (col < 0)
            continue;
        const int dsr = row * rb.src_subrows + subrow;
        const int dsc = col * rb.src_subcols + subcol;
        const int drow = dsr / rb.dest_subrows;
        const int dcol = dsc / rb.dest_subcols;
        const int b = find_block(d_offsets, d_row_counts, d_columns, drow, dcol);
        if (b < 0)
            continue;
        const int rps = src_br / rb.src_subrows;
        const int cps = src_bc / rb.src_subcols;
        const int di0 = (dsr - rb.dest_subrows * drow) * rps;
        const int dj0 = (dsc - rb.dest_subcols * dcol) * cps;
        const int si0 = subrow * rps;
        const int sj0 = subcol * cps;
        const SS* s = src + (size_t)p * src_br * src_bc;
        DS* d = dst + (size_t)b * dst_br * dst_bc;
        for (int i = 0; i < rps; ++i)
            for (int j = 0; j < cps; ++j)
                d[(di0 + i) * dst_bc + dj0 + j] = (DS)(scale * s[(si0 + i) * src_bc + sj0 + j]);
    }
}

// Padded same-shape assignment, one wave per row: the row's active source
// blocks, in storage order, go to the start of the destination row's capacity
// (reference _bsr_assign_padded_row_ranges and _bsr_assign_padded_copy_blocks);
// a row that does not fit is left empty and raises *status.
template <typename SS, typename DS>
__global__ __launch_bounds__(kWave) void assign_padded_kernel(
    int nrow, int block_size, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const SS* __restrict__ src, SS scale, const int* __restrict__ d_offsets,
    int* __restrict__ d_row_counts, int* __restrict__ d_columns, DS* __restrict__ dst, int* __restrict__ status)
{
    const int r = blockIdx.x;
    if (r >= nrow)
        return;
    const int beg = offsets[r];
    const int cap_end = offsets[r + 1];
    int end = row_counts ? beg + row_counts[r] : cap_end;
    end = end < cap_end ? end : cap_end;
    const int cnt = end - beg;
    const int dbeg = d_offsets[r];
    if (cnt > d_offsets[r + 1] - dbeg) {
        if (threadIdx.x == 0) {
            d_row_counts[r] = 0;
            if (status)
                atomicMax(status, kStatusRowCapacityExceeded);
        }
        return;
    }
    if (threadIdx.x == 0)
        d_row_counts[r] = cnt;
    for (int j = threadIdx.x; j < cnt; j += kWave)
        d_columns[dbeg + j] = columns[beg + j];
    if (src)
        for (int64_t e = threadIdx.x; e < (int64_t)cnt * block_size; e += kWave)
            dst[(size_t)dbeg * block_size + e] = (DS)(scale * src[(size_t)beg * block_size + e]);
}

// First position in [lo, hi) whose column is >= v (columns sorted).
__device__ __forceinline__ int lower_bound_col(const int* __restrict__ columns, int lo, int hi, int v)
{
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (columns[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Padded topology of a reblocked copy, one thread per destination row
// (reference _bsr_assign_padded_reblock_topology and
// _bsr_reblock_row_next_col, warp/_src/sparse.py:2000-2083): the
// destination columns are the sorted union, over the source rows covering
// the destination row, of the destination columns each source block
// overlaps.  Each step takes the smallest destination column above the
// previous one: per source row, the first block at or past the source column
// holding that column's first sub-column (binary search) gives the row's
// candidate.  A row that does not fit its capacity is left empty and raises
// *status.
__global__ __launch_bounds__(kBlock) void reblock_topology_kernel(
    Reblock rb, int src_nrow, int dest_nrow, int dest_ncol, const int* __restrict__ s_off,
    const int* __restrict__ s_rc, const int* __restrict__ s_col, const int* __restrict__ d_off,
    int* __restrict__ d_rc, int* __restrict__ d_col, int* __restrict__ status)
{
    const int drow = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (drow >= dest_nrow)
        return;
    const int64_t first_sub = (int64_t)drow * rb.dest_subrows;
    const int r0 = (int)(first_sub / rb.src_subrows);
    const int r1 = (int)min<int64_t>(src_nrow, (first_sub + rb.dest_subrows + rb.src_subrows - 1) / rb.src_subrows);
    const int beg = d_off[drow];
    const int cap = d_off[drow + 1];
    int b = beg;
    int prev = -1;
    bool overflow = false;
    for (;;) {
        const int minc = prev + 1;
        int next = dest_ncol;
        for (int r = r0; r < r1; ++r) {
            const int sb = s_off[r];
            const int se = row_end(s_off, s_rc, r);
            const int sc = (int)(((int64_t)minc * rb.dest_subcols) / rb.src_subcols);
            const int k = lower_bound_col(s_col, sb, se, sc);
            if (k == se)
                continue;
            const int64_t sub0 = (int64_t)s_col[k] * rb.src_subcols;
            const int dbeg = (int)(sub0 / rb.dest_subcols);
            const int dend = (int)((sub0 + rb.src_subcols + rb.dest_subcols - 1) / rb.dest_subcols);
            const int dc = max(dbeg, minc);
            if (dc < dend && dc < next)
                next = dc;
        }
        if (next >= dest_ncol)
            break;
        if (b >= cap) {
            overflow = true;
            break;
        }
        d_col[b++] = next;
        prev = next;
    }
    if (overflow) {
        d_rc[drow] = 0;
        if (status)
            atomicMax(status, kStatusRowCapacityExceeded);
    } else {
        d_rc[drow] = b - beg;
    }
}

// Row capacity of a reblocked padded copy (reference
// _bsr_copy_reblocked_capacity_counts, warp/_src/sparse.py:2086-2108): the
// source capacity of the covering source rows times the sub-columns per
// source block; cap[drow] (an exclusive scan turns it into offsets).
__global__ __launch_bounds__(kBlock) void reblock_capacity_kernel(Reblock rb, int src_nrow, int dest_nrow,
                                                                  const int* __restrict__ s_off, int* __restrict__ cap)
{
    const int drow = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (drow >= dest_nrow)
        return;
    const int64_t first_sub = (int64_t)drow * rb.dest_subrows;
    const int r0 = (int)(first_sub / rb.src_subrows);
    const int r1 = (int)min<int64_t>(src_nrow, (first_sub + rb.dest_subrows + rb.src_subrows - 1) / rb.src_subrows);
    int64_t n = 0;
    for (int r = r0; r < r1; ++r)
        n += (int64_t)(s_off[r + 1] - s_off[r]) * rb.src_subcols;
    cap[drow] = (int)n;
}

// rows[p] = r for the active positions p < T of row r (one wavefront per row;
// the other positions were set to -1 beforehand).
__global__ __launch_bounds__(kWave * kRowsPerBlock) void uncompress_rows_kernel(int nrow,
                                                                               const int* __restrict__ offsets,
                                                                               const int* __restrict__ row_counts,
                                                                               int T, int* __restrict__ rows)
{
    const int row = wave_row();
    if (row >= nrow)
        return;
    const int end = min(row_end(offsets, row_counts, row), T);
    for (int p = offsets[row] + (int)(threadIdx.x % kWave); p < end; p += kWave)
        rows[p] = row;
}

// dst[block (rows[i], cols[i])] += scale * src[i], one thread per source
// block (one search, then the block's scalars); blocks absent from the
// destination are skipped (reference _bsr_axpy_add_block).  Every destination
// block must receive at most one source block per launch.
template <typename S, int BS>
__global__ __launch_bounds__(kBlock) void add_blocks_kernel(int n, int block_size, const int* __restrict__ rows,
                                                           const int* __restrict__ cols, const S* __restrict__ src,
                                                           S scale, const int* __restrict__ d_offsets,
                                                           const int* __restrict__ d_row_counts,
                                                           const int* __restrict__ d_columns, S* __restrict__ dst)
{
#pragma clang fp contract(off)
    const int bs = BS > 0 ? BS : block_size;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int r = rows[i];
    if (r < 0)
        return;
    const int b = find_block(d_offsets, d_row_counts, d_columns, r, cols[i]);
    if (b < 0)
        return;
    const S* s = src + (size_t)i * bs;
    S* d = dst + (size_t)b * bs;
    if constexpr (BS > 0) {
#pragma unroll
        for (int e = 0; e < BS; ++e)
            d[e] += scale * s[e];
    } else {
        for (int e = 0; e < bs; ++e)
            d[e] += scale * s[e];
    }
}

// y := alpha x + beta y written into y's own row capacity, one wave per row
// (reference _bsr_axpy_padded_count / _fill / _finalize): the union of the
// two sorted rows is merged from the back, in place; a row whose union does
// not fit is left unchanged with row_counts 0 and raises *status.
template <typename S>
__global__ __launch_bounds__(kWave) void axpy_padded_kernel(
    int nrow, int bs, S alpha, S beta, const int* __restrict__ x_offsets, const int* __restrict__ x_row_counts,
    const int* __restrict__ x_columns, const S* __restrict__ x_values, const int* __restrict__ y_offsets,
    int* __restrict__ y_row_counts, int* __restrict__ y_columns, S* __restrict__ y_values, int* __restrict__ status)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x;
    if (r >= nrow)
        return;
    const int lane = threadIdx.x;
    const int xb = x_offsets[r];
    const int xe = row_end(x_offsets, x_row_counts, r);
    const int yb = y_offsets[r];
    const int ye = row_end(y_offsets, y_row_counts, r);
    int i = xb, j = yb, count = 0;
    while (i < xe && j < ye) {
        const int xc = x_columns[i];
        const int yc = y_columns[j];
        ++count;
        i += xc <= yc ? 1 : 0;
        j += yc <= xc ? 1 : 0;
    }
    count += (xe - i) + (ye - j);
    if (yb + count > y_offsets[r + 1]) {
        if (lane == 0) {
            y_row_counts[r] = 0;
            atomicMax(status, kStatusRowCapacityExceeded);
        }
        return;
    }
    i = xe - 1;
    j = ye - 1;
    for (int w = yb + count - 1; w >= yb; --w) {
        bool use_x = false, use_y = false;
        int col;
        if (i >= xb && j >= yb) {
            const int xc = x_columns[i];
            const int yc = y_columns[j];
            use_x = xc >= yc;
            use_y = yc >= xc;
            col = use_x ? xc : yc;
        } else if (i >= xb) {
            use_x = true;
            col = x_columns[i];
        } else {
            use_y = true;
            col = y_columns[j];
        }
        for (int e = lane; e < bs; e += kWave) {
            S v = S(0);
            if (use_x)
                v += alpha * x_values[(size_t)i * bs + e];
            if (use_y)
                v += beta * y_values[(size_t)j * bs + e];
            y_values[(size_t)w * bs + e] = v;
        }
        if (lane == 0)
            y_columns[w] = col;
        i -= use_x ? 1 : 0;
        j -= use_y ? 1 : 0;
    }
    if (lane == 0)
        y_row_counts[r] = count;
}

// Sparse product x @ y, symbolic count (reference bsr_mm_count_coeffs): for
// each active x block k of row r, counts[k] = number of active blocks in y
// row x_columns[k].  With deep_ranges, a row whose product count exceeds the
// span [min_col, max_col] of the columns it reaches is listed as that whole
// span instead: row_min[r] = min_col and the span size goes to the row's last
// x block (other blocks 0).  row_min[r] = -1 otherwise.  counts is zeroed
// beforehand.
__global__ __launch_bounds__(kBlock) void mm_count_kernel(int x_nrow, int y_nrow, int y_ncol, int deep_ranges,
                                                         const int* __restrict__ x_offsets,
                                                         const int* __restrict__ x_row_counts,
                                                         const int* __restrict__ x_columns,
                                                         const int* __restrict__ y_offsets,
                                                         const int* __restrict__ y_row_counts,
                                                         const int* __restrict__ y_columns, int* __restrict__ row_min,
                                                         int* __restrict__ counts)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= x_nrow)
        return;
    const int xb = x_offsets[r];
    const int xe = row_end(x_offsets, x_row_counts, r);
    int64_t total = 0;
    int min_col = y_ncol, max_col = 0;
    for (int k = xb; k < xe; ++k) {
        const int c = x_columns[k];
        int n = 0;
        if (c >= 0 && c < y_nrow) {
            const int b = y_offsets[c];
            const int e = row_end(y_offsets, y_row_counts, c);
            n = e - b;
            if (n > 0) {
                min_col = min(min_col, y_columns[b]);
                max_col = max(max_col, y_columns[e - 1]);
            }
        }
        counts[k] = n;
        total += n;
    }
    const int span = max(0, max_col - min_col + 1);
    if (deep_ranges && total > span) {
        row_min[r] = min_col;
        for (int k = xb; k < xe - 1; ++k)
            counts[k] = 0;
        counts[xe - 1] = span;
    } else {
        row_min[r] = -1;
    }
}

// Product list (reference _bsr_mm_list_coeffs): entry i < n of the product
// triplets, at base + i, is (row, column) of the product block numbered i in
// x-block order (mm_offsets = exclusive scan of the counts over the T x
// positions); entries at or beyond the real total get (-1, -1).
__global__ __launch_bounds__(kBlock) void mm_list_kernel(int n, int base, int T, int x_nrow,
                                                        const int* __restrict__ x_offsets,
                                                        const int* __restrict__ x_row_counts,
                                                        const int* __restrict__ x_columns,
                                                        const int* __restrict__ y_offsets,
                                                        const int* __restrict__ y_columns,
                                                        const int* __restrict__ row_min,
                                                        const int* __restrict__ mm_offsets, int* __restrict__ rows,
                                                        int* __restrict__ cols)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    if (i >= mm_offsets[T]) {
        rows[base + i] = -1;
        cols[base + i] = -1;
        return;
    }
    // last x position k with mm_offsets[k] <= i
    int lo = 0, hi = T - 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo + 1) / 2;
        if (mm_offsets[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int k = lo;
    const int pos = i - mm_offsets[k];
    const int row = row_of_position(x_offsets, x_row_counts, x_nrow, k);
    const int rmin = row_min[row];
    rows[base + i] = row;
    cols[base + i] = rmin < 0 ? y_columns[y_offsets[x_columns[k]] + pos] : rmin + pos;
}

// Product values (reference _bsr_mm_compute_values): for every active z
// block (r, c), acc = sum over the active x blocks k of row r, in storage
// order, of x_k @ y(x_columns[k], c) for those y blocks that exist; then
// z = (0 + beta * z) + alpha * acc (beta 1 leaves z's bits, -0 aside).  One thread per z block.  Each block product entry is
// x[i][0] y[0][j] + x[i][1] y[1][j] + ... evaluated left to right with every
// product and sum rounded (no fused multiply-add), so that the CPU oracle
// reproduces the bits.
template <typename S, int BR, int BK, int BC>
__global__ __launch_bounds__(kBlock) void mm_values_kernel(
    int T, int z_nrow, S alpha, S beta, const int* __restrict__ z_offsets, const int* __restrict__ z_row_counts,
    const int* __restrict__ z_columns, S* __restrict__ z_values, int y_nrow, const int* __restrict__ x_offsets,
    const int* __restrict__ x_row_counts, const int* __restrict__ x_columns, const S* __restrict__ x_values,
    const int* __restrict__ y_offsets, const int* __restrict__ y_row_counts, const int* __restrict__ y_columns,
    const S* __restrict__ y_values)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= T)
        return;
    const int r = row_of_position(z_offsets, z_row_counts, z_nrow, p);
    if (r < 0)
        return;
    const int col = z_columns[p];
    S acc[BR * BC];
#pragma unroll
    for (int e = 0; e < BR * BC; ++e)
        acc[e] = S(0);
    const int xb = x_offsets[r];
    const int xe = row_end(x_offsets, x_row_counts, r);
    for (int k = xb; k < xe; ++k) {
        const int xc = x_columns[k];
        if (xc < 0 || xc >= y_nrow)
            continue;
        const int b = find_block(y_offsets, y_row_counts, y_columns, xc, col);
        if (b < 0)
            continue;
        const S* xv = x_values + (size_t)k * BR * BK;
        const S* yv = y_values + (size_t)b * BK * BC;
#pragma unroll
        for (int i = 0; i < BR; ++i)
#pragma unroll
            for (int j = 0; j < BC; ++j) {
                S prod = xv[i * BK] * yv[j];
#pragma unroll
                for (int m = 1; m < BK; ++m)
                    prod = prod + xv[i * BK + m] * yv[m * BC + j];
                acc[i * BC + j] = acc[i * BC + j] + prod;
            }
    }
    S* zv = z_values + (size_t)p * BR * BC;
#pragma unroll
    for (int e = 0; e < BR * BC; ++e)
        zv[e] = (S(0) + beta * zv[e]) + alpha * acc[e];
}

// Tiled product values (replaces the reference's tiled-outer values kernel,
// make_bsr_mm_compute_values_tiled_outer, warp/_src/sparse.py:3813-3909):
// G lanes per z block.  The block's x row is read as the flat list of
// (x block k, inner index m) pairs, m fastest; lane g takes pairs g, g + G,
// ... and adds the rank-one update x_k[:, m] y(x_col k, c)[m, :] (each product
// rounded, then added), skipping pairs whose y block is absent; the G
// partial blocks are then combined by an xor butterfly (lane g adds lane
// g ^ off for off = G/2 .. 1).  Finally z = (0 + beta z) + alpha acc.  Block
// entries br * bc <= MAXE.
template <typename S, int MAXE>
__global__ __launch_bounds__(kBlock) void mm_values_tiled_kernel(
    int T, int G, int z_nrow, S alpha, S beta, int BR, int BK, int BC, const int* __restrict__ z_offsets,
    const int* __restrict__ z_row_counts, const int* __restrict__ z_columns, S* __restrict__ z_values, int y_nrow,
    const int* __restrict__ x_offsets, const int* __restrict__ x_row_counts, const int* __restrict__ x_columns,
    const S* __restrict__ x_values, const int* __restrict__ y_offsets, const int* __restrict__ y_row_counts,
    const int* __restrict__ y_columns, const S* __restrict__ y_values)
{
#pragma clang fp contract(off)
    const int p = (int)((blockIdx.x * kBlock + threadIdx.x) / G);
    const int lane = (int)(threadIdx.x % G);
    const int ne = BR * BC;
    int r = -1;
    int col = 0;
    if (p < T) {
        r = row_of_position(z_offsets, z_row_counts, z_nrow, p);
        col = r >= 0 ? z_columns[p] : 0;
    }
    S acc[MAXE];
#pragma unroll
    for (int e = 0; e < MAXE; ++e)
        acc[e] = S(0);
    if (r >= 0) {
        const int xb = x_offsets[r];
        const int n = (row_end(x_offsets, x_row_counts, r) - xb) * BK;
        for (int c = lane; c < n; c += G) {
            const int kk = c / BK;
            const int m = c - kk * BK;
            const int k = xb + kk;
            const int xc = x_columns[k];
            if (xc < 0 || xc >= y_nrow)
                continue;
            const int b = find_block(y_offsets, y_row_counts, y_columns, xc, col);
            if (b < 0)
                continue;
            const S* xv = x_values + (size_t)k * BR * BK;
            const S* yv = y_values + (size_t)b * BK * BC;
#pragma unroll
            for (int e = 0; e < MAXE; ++e)
                if (e < ne) {
                    const int i = e / BC;
                    const int j = e - i * BC;
                    acc[e] = acc[e] + xv[i * BK + m] * yv[m * BC + j];
                }
        }
    }
    // every lane of the wave takes part in the butterfly (uniform trip count)
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
        for (int e = 0; e < MAXE; ++e)
            acc[e] = acc[e] + __shfl_xor(acc[e], off, kWave);
    if (r >= 0 && lane == 0) {
        S* zv = z_values + (size_t)p * ne;
#pragma unroll
        for (int e = 0; e < MAXE; ++e)
            if (e < ne)
                zv[e] = (S(0) + beta * zv[e]) + alpha * acc[e];
    }
}

// Same, with block dimensions known at run time (br * bc <= 64).
template <typename S>
__global__ __launch_bounds__(kBlock) void mm_values_any_kernel(
    int T, int z_nrow, S alpha, S beta, int BR, int BK, int BC, const int* __restrict__ z_offsets,
    const int* __restrict__ z_row_counts, const int* __restrict__ z_columns, S* __restrict__ z_values, int y_nrow,
    const int* __restrict__ x_offsets, const int* __restrict__ x_row_counts, const int* __restrict__ x_columns,
    const S* __restrict__ x_values, const int* __restrict__ y_offsets, const int* __restrict__ y_row_counts,
    const int* __restrict__ y_columns, const S* __restrict__ y_values)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= T)
        return;
    const int r = row_of_position(z_offsets, z_row_counts, z_nrow, p);
    if (r < 0)
        return;
    const int col = z_columns[p];
    S acc[64];
    for (int e = 0; e < BR * BC; ++e)
        acc[e] = S(0);
    const int xb = x_offsets[r];
    const int xe = row_end(x_offsets, x_row_counts, r);
    for (int k = xb; k < xe; ++k) {
        const int xc = x_columns[k];
        if (xc < 0 || xc >= y_nrow)
            continue;
        const int b = find_block(y_offsets, y_row_counts, y_columns, xc, col);
        if (b < 0)
            continue;
        const S* xv = x_values + (size_t)k * BR * BK;
        const S* yv = y_values + (size_t)b * BK * BC;
        for (int i = 0; i < BR; ++i)
            for (int j = 0; j < BC; ++j) {
                S prod = xv[i * BK] * yv[j];
                for (int m = 1; m < BK; ++m)
                    prod = prod + xv[i * BK + m] * yv[m * BC + j];
                acc[i * BC + j] = acc[i * BC + j] + prod;
            }
    }
    S* zv = z_values + (size_t)p * BR * BC;
    for (int e = 0; e < BR * BC; ++e)
        zv[e] = (S(0) + beta * zv[e]) + alpha * acc[e];
}

// Block-diagonal topology: row r < ndiag holds column r.
__global__ __launch_bounds__(kBlock) void diag_topology_kernel(int nrow, int ndiag, int* __restrict__ offsets,
                                                              int* __restrict__ columns)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r > nrow)
        return;
    offsets[r] = r < ndiag ? r : ndiag;
    if (r < ndiag)
        columns[r] = r;
}

// y = alpha x + beta y (reference fem array_axpy), contraction off.
template <typename S>
__global__ __launch_bounds__(kBlock) void axpby_kernel(int64_t n, S alpha, const S* __restrict__ x, S beta,
                                                      S* __restrict__ y)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        y[i] = alpha * x[i] + beta * y[i];
}

constexpr int kMaxEigen = 8;

// Eigen decomposition of a symmetric n x n matrix (n <= kMaxEigen) by cyclic
// Jacobi rotations in FP64: on return a holds the eigenvalues on its diagonal
// and row k of v the unit eigenvector of eigenvalue a[k][k].
__device__ void symmetric_eigen(int n, double (&a)[kMaxEigen][kMaxEigen], double (&v)[kMaxEigen][kMaxEigen])
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            v[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += a[i][i] * a[i][i];
            for (int j = i + 1; j < n; ++j)
                off += a[i][j] * a[i][j];
        }
        if (off <= 1e-30 * diag || off == 0.0)
            break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p][q];
                if (apq == 0.0)
                    continue;
                // rotation angle zeroing a[p][q] (Golub & Van Loan, sym.schur2)
                const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - sn * akq;
                    a[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - sn * aqk;
                    a[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vpk = v[p][k], vqk = v[q][k];
                    v[p][k] = c * vpk - sn * vqk;
                    v[q][k] = sn * vpk + c * vqk;
                }
            }
    }
}

// Dirichlet projector normalization, one thread per block row (reference
// fem _normalize_dirichlet_projector(_and_values)_kernel and
// _normalize_projector_and_value, warp/_src/fem/dirichlet.py:127-199): the
// diagonal block P becomes the orthogonal projector sum v v^T over its
// eigenpairs with eigenvalue > eps = 1e-6 trace(P), and the fixed value b
// (when given) becomes sum (v . b) / lambda v over the same pairs.
template <typename S>
__global__ __launch_bounds__(kBlock) void normalize_projector_kernel(int nrow, int n, const int* __restrict__ offsets,
                                                                    const int* __restrict__ row_counts,
                                                                    const int* __restrict__ columns,
                                                                    S* __restrict__ values, S* __restrict__ fixed)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= nrow)
        return;
    const int b = find_block(offsets, row_counts, columns, r, r);
    if (b < 0)
        return;
    S* P = values + (size_t)b * n * n;
    double a[kMaxEigen][kMaxEigen], v[kMaxEigen][kMaxEigen];
    double trace = 0.0;
    for (int i = 0; i < n; ++i) {
        trace += (double)P[i * n + i];
        for (int j = 0; j < n; ++j)
            a[i][j] = 0.5 * ((double)P[i * n + j] + (double)P[j * n + i]);
    }
    const double eps = trace * 1.0e-6;
    symmetric_eigen(n, a, v);
    double out[kMaxEigen][kMaxEigen], bn[kMaxEigen];
    for (int i = 0; i < n; ++i) {
        bn[i] = 0.0;
        for (int j = 0; j < n; ++j)
            out[i][j] = 0.0;
    }
    for (int k = 0; k < n; ++k) {
        const double lam = a[k][k];
        if (!(lam > eps))
            continue;
        double vb = 0.0;
        if (fixed)
            for (int i = 0; i < n; ++i)
                vb += v[k][i] * (double)fixed[(size_t)r * n + i];
        for (int i = 0; i < n; ++i) {
            bn[i] += vb / lam * v[k][i];
            for (int j = 0; j < n; ++j)
                out[i][j] += v[k][i] * v[k][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            P[i * n + j] = (S)out[i][j];
    if (fixed)
        for (int i = 0; i < n; ++i)
            fixed[(size_t)r * n + i] = (S)bn[i];
}

template <typename S> __global__ __launch_bounds__(kBlock) void scale_kernel(int64_t n, S* __restrict__ x, S alpha)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        x[i] *= alpha;
}

// dst[i] = (DS)(scale * src[i]): the same-shape block copy of an assignment
// whose destination topology is the source's own.
template <typename SS, typename DS>
__global__ __launch_bounds__(kBlock) void copy_scaled_kernel(int64_t n, const SS* __restrict__ src, SS scale,
                                                            DS* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        dst[i] = (DS)(scale * src[i]);
}

// x = 0 + beta * x: the bits a union-topology rebuild gives a kept block
// (zeroed, then beta * old added), -0 included.
template <typename S> __global__ __launch_bounds__(kBlock) void rescale_kernel(int64_t n, S* __restrict__ x, S beta)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const S zero = S(0);
        x[i] = zero + beta * x[i];
    }
}

// Number of pairs (rows[i], cols[i]), rows[i] >= 0, with no block in the
// matrix, added to *missing (one atomic per workgroup).
__global__ __launch_bounds__(kBlock) void count_missing_kernel(int64_t n, const int* __restrict__ rows,
                                                              const int* __restrict__ cols,
                                                              const int* __restrict__ offsets,
                                                              const int* __restrict__ row_counts,
                                                              const int* __restrict__ columns, int* __restrict__ missing)
{
    __shared__ int lds_count;
    if (threadIdx.x == 0)
        lds_count = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n && rows[i] >= 0 && find_block(offsets, row_counts, columns, rows[i], cols[i]) < 0)
        atomicAdd(&lds_count, 1);
    __syncthreads();
    if (threadIdx.x == 0 && lds_count)
        atomicAdd(missing, lds_count);
}

template <typename SS, typename DS>
void launch_copy(unsigned grid, hipStream_t s, int T, Reblock rb, int nrow, const int* offsets, const int* row_counts,
                 const int* columns, int src_br, int src_bc, const void* src, double scale, const int* d_offsets,
                 const int* d_row_counts, const int* d_columns, int dst_br, int dst_bc, void* dst)
{
    assign_copy_blocks_kernel<SS, DS><<<grid, kWave * kRowsPerBlock, 0, s>>>(
        T, rb, nrow, offsets, row_counts, columns, src_br, src_bc, static_cast<const SS*>(src), (SS)scale, d_offsets,
        d_row_counts, d_columns, dst_br, dst_bc, static_cast<DS*>(dst));
}

}  // namespace
}  // namespace wa

// Destination block coordinates of every source sub-block for an assignment
// that splits source blocks into src_subrows x src_subcols pieces, each landing
// in the destination block that groups dest_subrows x dest_subcols pieces
// (reference _bsr_assign_list_blocks).  Output index
// (p * src_subcols + subcol) * src_subrows + subrow for source position p < T;
// (-1, -1) for positions that hold no active block.
WA_API int warp_amd_bsr_assign_list_blocks(int T, int src_subrows, int src_subcols, int dest_subrows, int dest_subcols,
                                           int nrow, const int* offsets, const int* row_counts, const int* columns,
                                           int* dest_rows, int* dest_cols, void* stream)
{
    WA_REQUIRE(T >= 0 && nrow >= 0 && src_subrows > 0 && src_subcols > 0 && dest_subrows > 0 && dest_subcols > 0,
               "bsr_assign: invalid sizes");
    const int64_t n = (int64_t)T * src_subrows * src_subcols;
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(offsets && columns && dest_rows && dest_cols, "bsr_assign: null arrays");
    hipStream_t s = wa::as_stream(stream);
    WA_TRY(wa::launch_fill_i32(dest_rows, n, -1, s));
    WA_TRY(wa::launch_fill_i32(dest_cols, n, -1, s));
    if (nrow == 0)
        return wa::kOk;
    wa::assign_list_blocks_kernel<<<(unsigned)wa::ceil_div(nrow, wa::kRowsPerBlock), wa::kWave * wa::kRowsPerBlock, 0,
                                    s>>>(T, wa::Reblock{src_subrows, src_subcols, dest_subrows, dest_subcols}, nrow,
                                         offsets, row_counts, columns, dest_rows, dest_cols);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Copy the values of every active source block (positions < T) into the
// destination topology, scaled in the source type then converted (reference
// _bsr_assign_copy_blocks).  Sub-blocks without a destination block are
// skipped; destination values are not cleared here.  Scalar sizes 4 or 8.
WA_API int warp_amd_bsr_assign_copy_blocks(int src_scalar_size, int dest_scalar_size, int T, int src_subrows,
                                           int src_subcols, int dest_subrows, int dest_subcols, int nrow,
                                           const int* offsets, const int* row_counts, const int* columns, int src_br,
                                           int src_bc, const void* src_values, double scale, const int* dest_offsets,
                                           const int* dest_row_counts, const int* dest_columns, int dest_br,
                                           int dest_bc, void* dest_values, void* stream)
{
    WA_REQUIRE(T >= 0 && nrow >= 0 && src_subrows > 0 && src_subcols > 0 && dest_subrows > 0 && dest_subcols > 0 &&
                   src_br > 0 && src_bc > 0 && dest_br > 0 && dest_bc > 0,
               "bsr_assign: invalid sizes");
    WA_REQUIRE((src_scalar_size == 4 || src_scalar_size == 8) && (dest_scalar_size == 4 || dest_scalar_size == 8),
               "bsr_assign: values must be fp32 or fp64");
    WA_REQUIRE(src_br % src_subrows == 0 && src_bc % src_subcols == 0 &&
                   src_br / src_subrows * dest_subrows == dest_br && src_bc / src_subcols * dest_subcols == dest_bc,
               "bsr_assign: incompatible block shapes");
    const int64_t n = (int64_t)T * src_subrows * src_subcols;
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(offsets && columns && src_values && dest_offsets && dest_values, "bsr_assign: null arrays");
    if (nrow == 0)
        return wa::kOk;
    const unsigned grid = (unsigned)wa::ceil_div(nrow, wa::kRowsPerBlock);
    const wa::Reblock rb{src_subrows, src_subcols, dest_subrows, dest_subcols};
    hipStream_t s = wa::as_stream(stream);
#define WA_COPY(SS, DS)                                                                                               \
    wa::launch_copy<SS, DS>(grid, s, T, rb, nrow, offsets, row_counts, columns, src_br, src_bc, src_values, scale,    \
                            dest_offsets, dest_row_counts, dest_columns, dest_br, dest_bc, dest_values)
    if (src_scalar_size == 4 && dest_scalar_size == 4)
        WA_COPY(float, float);
    else if (src_scalar_size == 4)
        WA_COPY(float, double);
    else if (dest_scalar_size == 4)
        WA_COPY(double, float);
    else
        WA_COPY(double, double);
#undef WA_COPY
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Padded assignment of a matrix into a destination with the same block shape
// (reference bsr_assign with topology="padded", warp/_src/sparse.py:2331-2368):
// row r's active source blocks, in storage order, are copied to
// d_offsets[r].. and d_row_counts[r] set; values (src_values NULL: structure
// only) are scaled in the source type then converted.  A row exceeding its
// capacity is left empty and raises *status (sticky) to 1.
WA_API int warp_amd_bsr_assign_padded(int src_scalar_size, int dest_scalar_size, int nrow, int block_size,
                                      const int* offsets, const int* row_counts, const int* columns,
                                      const void* src_values, double scale, const int* dest_offsets,
                                      int* dest_row_counts, int* dest_columns, void* dest_values, int* status,
                                      void* stream)
{
    WA_REQUIRE(nrow >= 0 && block_size > 0, "bsr_assign: invalid sizes");
    WA_REQUIRE(!src_values || ((src_scalar_size == 4 || src_scalar_size == 8) &&
                               (dest_scalar_size == 4 || dest_scalar_size == 8) && dest_values),
               "bsr_assign: values must be fp32 or fp64");
    if (nrow == 0)
        return wa::kOk;
    WA_REQUIRE(offsets && dest_offsets && dest_row_counts, "bsr_assign: null arrays");
    hipStream_t s = wa::as_stream(stream);
    const unsigned grid = (unsigned)nrow;
#define WA_PAD(SS, DS)                                                                                                \
    wa::assign_padded_kernel<SS, DS><<<grid, wa::kWave, 0, s>>>(                                                      \
        nrow, block_size, offsets, row_counts, columns, static_cast<const SS*>(src_values), (SS)scale, dest_offsets,  \
        dest_row_counts, dest_columns, static_cast<DS*>(dest_values), status)
    if (!src_values || (src_scalar_size == 4 && dest_scalar_size == 4))
        WA_PAD(float, float);
    else if (src_scalar_size == 4)
        WA_PAD(float, double);
    else if (dest_scalar_size == 4)
        WA_PAD(double, float);
    else
        WA_PAD(double, double);
#undef WA_PAD
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Padded topology of a reblocked assignment (reference bsr_assign with
// topology="padded" and a change of block shape, warp/_src/sparse.py:
// 2201-2255, 2346-2358): dest_columns / dest_row_counts of each destination
// row within its existing capacity (dest_offsets); rows that do not fit are
// left empty and raise *status.  The values are then copied by
// warp_amd_bsr_assign_copy_blocks into the zeroed destination.
WA_API int warp_amd_bsr_assign_padded_reblock_topology(int src_subrows, int src_subcols, int dest_subrows,
                                                       int dest_subcols, int src_nrow, int dest_nrow, int dest_ncol,
                                                       const int* src_offsets, const int* src_row_counts,
                                                       const int* src_columns, const int* dest_offsets,
                                                       int* dest_row_counts, int* dest_columns, int* status,
                                                       void* stream)
{
    WA_REQUIRE(src_subrows > 0 && src_subcols > 0 && dest_subrows > 0 && dest_subcols > 0 && src_nrow >= 0 &&
                   dest_nrow >= 0 && dest_ncol >= 0,
               "bsr_assign: invalid reblocking sizes");
    if (dest_nrow == 0)
        return wa::kOk;
    WA_REQUIRE(src_offsets && dest_offsets && dest_row_counts && (src_nrow == 0 || src_columns),
               "bsr_assign: null arrays");
    const wa::Reblock rb{src_subrows, src_subcols, dest_subrows, dest_subcols};
    wa::reblock_topology_kernel<<<(unsigned)wa::ceil_div(dest_nrow, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        rb, src_nrow, dest_nrow, dest_ncol, src_offsets, src_row_counts, src_columns, dest_offsets, dest_row_counts,
        dest_columns, status);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Workspace of warp_amd_bsr_reblock_capacity (its offsets scan).
WA_API size_t warp_amd_bsr_reblock_capacity_workspace_size(int dest_nrow)
{
    return wa::scan_workspace_bytes((int64_t)(dest_nrow > 0 ? dest_nrow : 0) + 1);
}

// Offsets of a reblocked padded copy (reference bsr_copy with
// topology="padded" and a change of block shape, warp/_src/sparse.py:
// 2561-2588): each destination row holds the capacity of its covering source
// rows times the sub-columns per source block; dest_offsets[dest_nrow + 1].
WA_API int warp_amd_bsr_reblock_capacity(int src_subrows, int src_subcols, int dest_subrows, int dest_subcols,
                                         int src_nrow, int dest_nrow, const int* src_offsets, int* dest_offsets,
                                         void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(src_subrows > 0 && src_subcols > 0 && dest_subrows > 0 && dest_subcols > 0 && src_nrow >= 0 &&
                   dest_nrow >= 0,
               "bsr_copy: invalid reblocking sizes");
    WA_REQUIRE(dest_offsets, "bsr_copy: null offsets");
    hipStream_t s = wa::as_stream(stream);
    if (dest_nrow == 0)
        return wa::launch_fill_i32(dest_offsets, 1, 0, s);
    WA_REQUIRE(src_offsets, "bsr_copy: null source offsets");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_bsr_reblock_capacity_workspace_size(dest_nrow),
               "bsr_copy: workspace too small");
    const wa::Reblock rb{src_subrows, src_subcols, dest_subrows, dest_subcols};
    wa::reblock_capacity_kernel<<<(unsigned)wa::ceil_div(dest_nrow, wa::kBlock), wa::kBlock, 0, s>>>(
        rb, src_nrow, dest_nrow, src_offsets, dest_offsets);
    WA_LAUNCH_CHECK();
    return wa::exclusive_scan_i32(dest_offsets, dest_offsets, dest_nrow, workspace, s);
}

// x[0..n) *= alpha (reference bsr_scale's value kernel); fp32 or fp64.
WA_API int warp_amd_copy_scaled_values(int src_scalar_size, int dst_scalar_size, int64_t n, const void* src,
                                       double scale, void* dst, void* stream)
{
    WA_REQUIRE(n >= 0, "copy_scaled_values: invalid size");
    WA_REQUIRE((src_scalar_size == 4 || src_scalar_size == 8) && (dst_scalar_size == 4 || dst_scalar_size == 8),
               "copy_scaled_values: values must be fp32 or fp64");
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(src && dst, "copy_scaled_values: null array");
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
#define WA_COPY(SS, DS)                                                                                               \
    wa::copy_scaled_kernel<SS, DS><<<grid, wa::kBlock, 0, s>>>(n, static_cast<const SS*>(src), (SS)scale,             \
                                                               static_cast<DS*>(dst))
    if (src_scalar_size == 4 && dst_scalar_size == 4)
        WA_COPY(float, float);
    else if (src_scalar_size == 4)
        WA_COPY(float, double);
    else if (dst_scalar_size == 4)
        WA_COPY(double, float);
    else
        WA_COPY(double, double);
#undef WA_COPY
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_rescale_values(int scalar_size, int64_t n, void* x, double beta, void* stream)
{
    WA_REQUIRE(n >= 0, "rescale_values: invalid size");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "rescale_values: values must be fp32 or fp64");
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(x != nullptr, "rescale_values: null array");
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
    if (scalar_size == 4)
        wa::rescale_kernel<float><<<grid, wa::kBlock, 0, s>>>(n, static_cast<float*>(x), (float)beta);
    else
        wa::rescale_kernel<double><<<grid, wa::kBlock, 0, s>>>(n, static_cast<double*>(x), beta);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_bsr_count_missing_blocks(int64_t n, const int* rows, const int* cols, const int* offsets,
                                             const int* row_counts, const int* columns, int* missing, void* stream)
{
    WA_REQUIRE(n >= 0 && missing, "count_missing_blocks: invalid arguments");
    hipStream_t s = wa::as_stream(stream);
    WA_TRY(wa::launch_fill_i32(missing, 1, 0, s));
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(rows && cols && offsets && columns, "count_missing_blocks: null array");
    wa::count_missing_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, s>>>(n, rows, cols, offsets,
                                                                                         row_counts, columns, missing);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_scale_values(int scalar_size, int64_t n, void* x, double alpha, void* stream)
{
    WA_REQUIRE(n >= 0, "scale_values: invalid size");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "scale_values: values must be fp32 or fp64");
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(x != nullptr, "scale_values: null array");
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
    if (scalar_size == 4)
        wa::scale_kernel<float><<<grid, wa::kBlock, 0, s>>>(n, static_cast<float*>(x), (float)alpha);
    else
        wa::scale_kernel<double><<<grid, wa::kBlock, 0, s>>>(n, static_cast<double*>(x), alpha);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// ---------------------------------------------------------------------------
// Sums and products of BSR matrices (bsr_axpy, bsr_mm)
// ---------------------------------------------------------------------------

// rows[p] = block row of storage position p < T, -1 where p holds no active
// block (replaces BsrMatrix.uncompress_rows' kernel _bsr_get_block_row,
// warp/_src/sparse.py:164-193).
WA_API int warp_amd_bsr_uncompress_rows(int nrow, const int* offsets, const int* row_counts, int T, int* rows,
                                        void* stream)
{
    WA_REQUIRE(nrow >= 0 && T >= 0, "uncompress_rows: invalid sizes");
    if (T == 0)
        return wa::kOk;
    WA_REQUIRE(offsets && rows, "uncompress_rows: null arrays");
    hipStream_t s = wa::as_stream(stream);
    WA_TRY(wa::launch_fill_i32(rows, T, -1, s));
    if (nrow == 0)
        return wa::kOk;
    wa::uncompress_rows_kernel<<<(unsigned)wa::ceil_div(nrow, wa::kRowsPerBlock), wa::kWave * wa::kRowsPerBlock, 0,
                                 s>>>(nrow, offsets, row_counts, T, rows);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// dst[block (rows[i], cols[i])] += scale * src[i] for i < n (rows[i] < 0 or
// blocks missing from the destination topology are skipped); replaces
// _bsr_axpy_add_block (warp/_src/sparse.py:3145-3163).  Each destination
// block may receive at most one source block per call.
WA_API int warp_amd_bsr_add_blocks(int scalar_size, int block_size, int64_t n, const int* rows, const int* cols,
                                   const void* src, double scale, const int* dest_offsets,
                                   const int* dest_row_counts, const int* dest_columns, void* dest_values,
                                   void* stream)
{
    WA_REQUIRE(n >= 0 && n <= INT_MAX && block_size > 0, "bsr_add_blocks: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_add_blocks: values must be fp32 or fp64");
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(rows && cols && src && dest_offsets && dest_values, "bsr_add_blocks: null arrays");
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
#define WA_ADD(S, BS)                                                                                                 \
    wa::add_blocks_kernel<S, BS><<<grid, wa::kBlock, 0, s>>>((int)n, block_size, rows, cols,                          \
                                                            static_cast<const S*>(src), (S)scale, dest_offsets,      \
                                                            dest_row_counts, dest_columns, static_cast<S*>(dest_values))
    if (scalar_size == 4) {
        if (block_size == 9)
            WA_ADD(float, 9);
        else if (block_size == 1)
            WA_ADD(float, 1);
        else
            WA_ADD(float, 0);
    } else {
        if (block_size == 9)
            WA_ADD(double, 9);
        else if (block_size == 1)
            WA_ADD(double, 1);
        else
            WA_ADD(double, 0);
    }
#undef WA_ADD
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// y := alpha x + beta y into y's existing row capacity (bsr_axpy with
// topology="padded", warp/_src/sparse.py:3395-3478): per row, the sorted union
// of x's and y's active columns, values alpha x + beta y; a row whose union
// exceeds its capacity is left unchanged with row_counts 0 and raises *status.
// x and y must not alias.
WA_API int warp_amd_bsr_axpy_padded(int scalar_size, int block_size, int nrow, double alpha, double beta,
                                    const int* x_offsets, const int* x_row_counts, const int* x_columns,
                                    const void* x_values, const int* y_offsets, int* y_row_counts, int* y_columns,
                                    void* y_values, int* status, void* stream)
{
    WA_REQUIRE(nrow >= 0 && block_size > 0, "bsr_axpy: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_axpy: values must be fp32 or fp64");
    if (nrow == 0)
        return wa::kOk;
    WA_REQUIRE(x_offsets && y_offsets && y_row_counts && x_values && y_values && status, "bsr_axpy: null arrays");
    hipStream_t s = wa::as_stream(stream);
    if (scalar_size == 4)
        wa::axpy_padded_kernel<float><<<(unsigned)nrow, wa::kWave, 0, s>>>(
            nrow, block_size, (float)alpha, (float)beta, x_offsets, x_row_counts, x_columns,
            static_cast<const float*>(x_values), y_offsets, y_row_counts, y_columns, static_cast<float*>(y_values),
            status);
    else
        wa::axpy_padded_kernel<double><<<(unsigned)nrow, wa::kWave, 0, s>>>(
            nrow, block_size, alpha, beta, x_offsets, x_row_counts, x_columns, static_cast<const double*>(x_values),
            y_offsets, y_row_counts, y_columns, static_cast<double*>(y_values), status);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Workspace of warp_amd_bsr_mm_count for T x positions.
WA_API size_t warp_amd_bsr_mm_count_workspace_size(int T)
{
    return wa::align_up((size_t)T * sizeof(int), 256) + wa::scan_workspace_bytes((int64_t)T + 1) + 256;
}

// Symbolic product x @ y (replaces bsr_mm_count_coeffs and its scan,
// warp/_src/sparse.py:3620-3680, 4455-4475): mm_offsets[0..T] = exclusive
// scan over the T x positions of the number of product blocks each lists
// (see deep_ranges below), row_min[0..x_nrow) as in the reference, and
// *total_host (pinned, optional, asynchronous) = mm_offsets[T].  deep_ranges =
// 1 lists a row reaching fewer distinct columns than it has products as its
// whole column span (the reference's compact-topology rule); 0 lists every
// product.
WA_API int warp_amd_bsr_mm_count(int T, int x_nrow, int y_nrow, int y_ncol, int deep_ranges, const int* x_offsets,
                                 const int* x_row_counts, const int* x_columns, const int* y_offsets,
                                 const int* y_row_counts, const int* y_columns, int* row_min, int* mm_offsets,
                                 int* total_host, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(T >= 0 && x_nrow >= 0 && y_nrow >= 0 && y_ncol >= 0, "bsr_mm: invalid sizes");
    WA_REQUIRE(x_offsets && y_offsets && row_min && mm_offsets, "bsr_mm: null arrays");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_bsr_mm_count_workspace_size(T) - 256,
               "bsr_mm: workspace too small");
    hipStream_t s = wa::as_stream(stream);
    int* counts = static_cast<int*>(workspace);
    void* scan_ws = static_cast<char*>(workspace) + wa::align_up((size_t)T * sizeof(int), 256);
    WA_TRY(wa::launch_fill_i32(counts, T, 0, s));
    if (x_nrow > 0) {
        wa::mm_count_kernel<<<(unsigned)wa::ceil_div(x_nrow, wa::kBlock), wa::kBlock, 0, s>>>(
            x_nrow, y_nrow, y_ncol, deep_ranges, x_offsets, x_row_counts, x_columns, y_offsets, y_row_counts,
            y_columns, row_min, counts);
        WA_LAUNCH_CHECK();
    }
    WA_TRY(wa::exclusive_scan_i32(counts, mm_offsets, T, scan_ws, s));
    if (total_host)
        WA_HIP(hipMemcpyAsync(total_host, mm_offsets + T, sizeof(int), hipMemcpyDeviceToHost, s));
    return wa::kOk;
}

// Product triplets (replaces _bsr_mm_list_coeffs, warp/_src/sparse.py:3683-3735):
// rows/cols[base + i] for i < n = (row, column) of product block i (in x-block
// order, from warp_amd_bsr_mm_count's mm_offsets / row_min); (-1, -1) for
// i >= mm_offsets[T].
WA_API int warp_amd_bsr_mm_list(int n, int base, int T, int x_nrow, const int* x_offsets, const int* x_row_counts,
                                const int* x_columns, const int* y_offsets, const int* y_columns, const int* row_min,
                                const int* mm_offsets, int* rows, int* cols, void* stream)
{
    WA_REQUIRE(n >= 0 && base >= 0 && T >= 0 && x_nrow >= 0, "bsr_mm: invalid sizes");
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(x_offsets && y_offsets && row_min && mm_offsets && rows && cols, "bsr_mm: null arrays");
    wa::mm_list_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, base, T, x_nrow, x_offsets, x_row_counts, x_columns, y_offsets, y_columns, row_min, mm_offsets, rows, cols);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Product values into an existing topology (replaces _bsr_mm_compute_values,
// warp/_src/sparse.py:3764-3800): for each active z block (r, c) among the
// first T positions, z += alpha * sum over x row r's active blocks k (storage
// order) of x_k @ y(x_columns[k], c).  Blocks: x br x bk, y bk x bc, z br x bc
// (br * bc <= 64); fp32 or fp64.  No fused multiply-add: products and sums
// are rounded one by one, in the order stated.  The old block enters as
// 0 + beta * z (beta 1: a topology rebuild already scaled it).
// Tiled product values: `tile` lanes per z block (a power of two in
// [1, 64]), rank-one updates per (x block, inner index) pair and an xor
// butterfly (replaces bsr_mm's tiled-outer values kernel, warp/_src/
// sparse.py:3813-3909, selected by tile_size > 0, 4101-4144); otherwise as
// warp_amd_bsr_mm_values.
WA_API int warp_amd_bsr_mm_values_tiled(int scalar_size, int br, int bk, int bc, double alpha, double beta, int T,
                                        int tile, int z_nrow, const int* z_offsets, const int* z_row_counts,
                                        const int* z_columns, void* z_values, int y_nrow, const int* x_offsets,
                                        const int* x_row_counts, const int* x_columns, const void* x_values,
                                        const int* y_offsets, const int* y_row_counts, const int* y_columns,
                                        const void* y_values, void* stream)
{
    WA_REQUIRE(br > 0 && bk > 0 && bc > 0 && br * bc <= 64 && T >= 0 && z_nrow >= 0, "bsr_mm: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_mm: values must be fp32 or fp64");
    WA_REQUIRE(tile >= 1 && tile <= wa::kWave && (tile & (tile - 1)) == 0,
               "bsr_mm: tile_size must be a power of two in [1, 64], got %d", tile);
    if (T == 0 || z_nrow == 0)
        return wa::kOk;
    WA_REQUIRE(z_offsets && z_columns && z_values && x_offsets && y_offsets, "bsr_mm: null arrays");
    const unsigned grid = (unsigned)wa::ceil_div((int64_t)T * tile, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
#define WA_MMT(S, MAXE)                                                                                               \
    wa::mm_values_tiled_kernel<S, MAXE><<<grid, wa::kBlock, 0, s>>>(                                                  \
        T, tile, z_nrow, (S)alpha, (S)beta, br, bk, bc, z_offsets, z_row_counts, z_columns, static_cast<S*>(z_values), \
        y_nrow, x_offsets, x_row_counts, x_columns, static_cast<const S*>(x_values), y_offsets, y_row_counts,          \
        y_columns, static_cast<const S*>(y_values))
    const bool small = br * bc <= 9;
    if (scalar_size == 4) {
        if (small)
            WA_MMT(float, 9);
        else
            WA_MMT(float, 64);
    } else {
        if (small)
            WA_MMT(double, 9);
        else
            WA_MMT(double, 64);
    }
#undef WA_MMT
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_bsr_mm_values(int scalar_size, int br, int bk, int bc, double alpha, double beta, int T,
                                  int z_nrow,
                                  const int* z_offsets, const int* z_row_counts, const int* z_columns,
                                  void* z_values, int y_nrow, const int* x_offsets, const int* x_row_counts,
                                  const int* x_columns, const void* x_values, const int* y_offsets,
                                  const int* y_row_counts, const int* y_columns, const void* y_values, void* stream)
{
    WA_REQUIRE(br > 0 && bk > 0 && bc > 0 && br * bc <= 64 && T >= 0 && z_nrow >= 0, "bsr_mm: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_mm: values must be fp32 or fp64");
    if (T == 0 || z_nrow == 0)
        return wa::kOk;
    WA_REQUIRE(z_offsets && z_columns && z_values && x_offsets && y_offsets, "bsr_mm: null arrays");
    const unsigned grid = (unsigned)wa::ceil_div(T, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
#define WA_MM(S, BR, BK, BC)                                                                                          \
    wa::mm_values_kernel<S, BR, BK, BC><<<grid, wa::kBlock, 0, s>>>(                               \
        T, z_nrow, (S)alpha, (S)beta, z_offsets, z_row_counts, z_columns, static_cast<S*>(z_values), y_nrow,         \
        x_offsets,                                                                                                    \
        x_row_counts, x_columns, static_cast<const S*>(x_values), y_offsets, y_row_counts, y_columns,                 \
        static_cast<const S*>(y_values))
#define WA_MM_ANY(S)                                                                                                  \
    wa::mm_values_any_kernel<S><<<grid, wa::kBlock, 0, s>>>(                                       \
        T, z_nrow, (S)alpha, (S)beta, br, bk, bc, z_offsets, z_row_counts, z_columns, static_cast<S*>(z_values),     \
        y_nrow,                                                                                                       \
        x_offsets, x_row_counts, x_columns, static_cast<const S*>(x_values), y_offsets, y_row_counts, y_columns,      \
        static_cast<const S*>(y_values))
    const bool s1 = br == 1 && bk == 1 && bc == 1;
    const bool s3 = br == 3 && bk == 3 && bc == 3;
    if (scalar_size == 4) {
        if (s1)
            WA_MM(float, 1, 1, 1);
        else if (s3)
            WA_MM(float, 3, 3, 3);
        else
            WA_MM_ANY(float);
    } else {
        if (s1)
            WA_MM(double, 1, 1, 1);
        else if (s3)
            WA_MM(double, 3, 3, 3);
        else
            WA_MM_ANY(double);
    }
#undef WA_MM
#undef WA_MM_ANY
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Block-diagonal topology of an nrow-row matrix: offsets[r] = min(r, ndiag),
// columns[r] = r for r < ndiag (replaces _bsr_set_diag_kernel of bsr_set_diag,
// warp/_src/sparse.py:2939-2995).
WA_API int warp_amd_bsr_diag_topology(int nrow, int ndiag, int* offsets, int* columns, void* stream)
{
    WA_REQUIRE(nrow >= 0 && ndiag >= 0 && ndiag <= nrow, "bsr_set_diag: invalid sizes");
    WA_REQUIRE(offsets && (ndiag == 0 || columns), "bsr_set_diag: null arrays");
    wa::diag_topology_kernel<<<(unsigned)wa::ceil_div((int64_t)nrow + 1, wa::kBlock), wa::kBlock, 0,
                               wa::as_stream(stream)>>>(nrow, ndiag, offsets, columns);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// y = alpha x + beta y over n scalars, fp32 or fp64 (reference fem array_axpy,
// warp/_src/fem/linalg.py:425-461).
WA_API int warp_amd_axpby(int scalar_size, int64_t n, double alpha, const void* x, double beta, void* y, void* stream)
{
    WA_REQUIRE(n >= 0, "axpby: invalid size");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "axpby: values must be fp32 or fp64");
    if (n == 0)
        return wa::kOk;
    WA_REQUIRE(x && y, "axpby: null arrays");
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
    if (scalar_size == 4)
        wa::axpby_kernel<float><<<grid, wa::kBlock, 0, s>>>(n, (float)alpha, static_cast<const float*>(x), (float)beta,
                                                            static_cast<float*>(y));
    else
        wa::axpby_kernel<double><<<grid, wa::kBlock, 0, s>>>(n, alpha, static_cast<const double*>(x), beta,
                                                             static_cast<double*>(y));
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Normalize the diagonal blocks (block_dim x block_dim, block_dim <= 8) of a
// Dirichlet projector into orthogonal projectors, and the fixed values
// (nrow x block_dim, may be NULL) accordingly (replaces
// normalize_dirichlet_projector's kernels, warp/_src/fem/dirichlet.py:12-66,
// 127-199).  The eigen decomposition runs in FP64 (Jacobi rotations) where the
// reference runs a QR iteration in the matrix scalar type.
WA_API int warp_amd_normalize_dirichlet_projector(int scalar_size, int block_dim, int nrow, const int* offsets,
                                                  const int* row_counts, const int* columns, void* values,
                                                  void* fixed_values, void* stream)
{
    WA_REQUIRE(nrow >= 0 && block_dim > 0 && block_dim <= wa::kMaxEigen, "normalize_dirichlet_projector: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "normalize_dirichlet_projector: values must be fp32 or fp64");
    if (nrow == 0)
        return wa::kOk;
    WA_REQUIRE(offsets && values, "normalize_dirichlet_projector: null arrays");
    const unsigned grid = (unsigned)wa::ceil_div(nrow, wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
    if (scalar_size == 4)
        wa::normalize_projector_kernel<float><<<grid, wa::kBlock, 0, s>>>(
            nrow, block_dim, offsets, row_counts, columns, static_cast<float*>(values), static_cast<float*>(fixed_values));
    else
        wa::normalize_projector_kernel<double><<<grid, wa::kBlock, 0, s>>>(
            nrow, block_dim, offsets, row_counts, columns, static_cast<double*>(values),
            static_cast<double*>(fixed_values));
    WA_LAUNCH_CHECK();
    return wa::kOk;
}
