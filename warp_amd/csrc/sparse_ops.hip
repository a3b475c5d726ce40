// Whole-matrix BSR operations around the solver: assignment with block
// reshaping and scalar conversion (bsr_assign / bsr_copy) and value scaling.
//
// Replaces the JIT kernels _bsr_assign_list_blocks and _bsr_assign_copy_blocks
// (warp/_src/sparse.py, launched from bsr_assign, 2258-2452) and
// _bsr_scale_kernel (bsr_scale).  The compact topology of an assignment is
// built by the triplet assembly (one triplet per destination sub-block).
#include "common.h"

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

struct Reblock {
    int src_subrows, src_subcols, dest_subrows, dest_subcols;
};

// One thread per (source position, sub-row, sub-column): the destination block
// coordinates of that sub-block, or (-1, -1).
__global__ __launch_bounds__(kBlock) void assign_list_blocks_kernel(int T, Reblock rb, int nrow,
                                                                   const int* __restrict__ offsets,
                                                                   const int* __restrict__ row_counts,
                                                                   const int* __restrict__ columns,
                                                                   int* __restrict__ dest_rows,
                                                                   int* __restrict__ dest_cols)
{
    const int per = rb.src_subrows * rb.src_subcols;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)T * per)
        return;
    const int p = (int)(t / per);
    const int rem = (int)(t - (int64_t)p * per);
    const int subrow = rem % rb.src_subrows;
    const int subcol = rem / rb.src_subrows;
    const int row = row_of_position(offsets, row_counts, nrow, p);
    const int col = row < 0 ? -1 : columns[p];
    if (row < 0 || col < 0) {
        dest_rows[t] = -1;
        dest_cols[t] = -1;
        return;
    }
    dest_rows[t] = (row * rb.src_subrows + subrow) / rb.dest_subrows;
    dest_cols[t] = (col * rb.src_subcols + subcol) / rb.dest_subcols;
}

// One thread per (source position, sub-row, sub-column): copies the sub-block
// into its destination block, converted to the destination scalar type after
// the scale is applied in the source type.
template <typename SS, typename DS>
__global__ __launch_bounds__(kBlock) void assign_copy_blocks_kernel(
    int T, Reblock rb, int nrow, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, int src_br, int src_bc, const SS* __restrict__ src, SS scale,
    const int* __restrict__ d_offsets, const int* __restrict__ d_row_counts, const int* __restrict__ d_columns,
    int dst_br, int dst_bc, DS* __restrict__ dst)
{
    const int per = rb.src_subrows * rb.src_subcols;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)T * per)
        return;
    const int p = (int)(t / per);
    const int rem = (int)(t - (int64_t)p * per);
    const int subrow = rem % rb.src_subrows;
    const int subcol = rem / rb.src_subrows;
    const int row = row_of_position(offsets, row_counts, nrow, p);
    if (row < 0)
        return;
    const int col = columns[p];
    if (col < 0)
        return;
    const int dsr = row * rb.src_subrows + subrow;
    const int dsc = col * rb.src_subcols + subcol;
    const int drow = dsr / rb.dest_subrows;
    const int dcol = dsc / rb.dest_subcols;
    const int b = find_block(d_offsets, d_row_counts, d_columns, drow, dcol);
    if (b < 0)
        return;
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

template <typename S> __global__ __launch_bounds__(kBlock) void scale_kernel(int64_t n, S* __restrict__ x, S alpha)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        x[i] *= alpha;
}

template <typename SS, typename DS>
void launch_copy(unsigned grid, hipStream_t s, int T, Reblock rb, int nrow, const int* offsets, const int* row_counts,
                 const int* columns, int src_br, int src_bc, const void* src, double scale, const int* d_offsets,
                 const int* d_row_counts, const int* d_columns, int dst_br, int dst_bc, void* dst)
{
    assign_copy_blocks_kernel<SS, DS><<<grid, kBlock, 0, s>>>(
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
    wa::assign_list_blocks_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        T, wa::Reblock{src_subrows, src_subcols, dest_subrows, dest_subcols}, nrow, offsets, row_counts, columns,
        dest_rows, dest_cols);
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
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
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

// x[0..n) *= alpha (reference bsr_scale's value kernel); fp32 or fp64.
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
