// bsr_mv, bsr_get_diag and Jacobi extraction for gfx950.
//
// Replaces the reference's JIT kernels bsr_mv_kernel / bsr_mv_tiled_kernel
// (warp/_src/sparse.py:4627-4715), _bsr_get_diag_kernel (sparse.py:2871-2884)
// and _extract_inverse_diagonal_* (warp/_src/optim/linear.py:2290-2320).
#include "bsr_core.h"

namespace wa {
namespace {

constexpr int kBlock = 256;

// Interleaved rows: a wave's kWave / G groups take its kWave / G * RPG rows
// round-robin (group g: rows g, g + kWave / G, ...), so the j-th block load of
// all groups reads consecutive rows' blocks.  Each row's sum is unchanged.
constexpr bool kInterleaveRows = true;

// RPG block rows per group of G lanes (see group_rows_product).
// reverse: the waves take their row ranges from the last to the first (each
// row's sum is unchanged).  Alternated between consecutive products with one
// matrix, the rows read last by one product are read first by the next, while
// they are still in the 256 MB memory-side cache (5.99 vs 5.13-5.38 TB/s on
// the 295 MB 64^3 elasticity product).
template <typename S, int BR, int BC, int G, int RPG>
__global__ __launch_bounds__(kBlock) void bsr_mv_rows_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ row_counts, const int* __restrict__ columns,
    const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y, S alpha, S beta, int reverse
)
{
    static_assert(G <= kWave, "a row group lies within one wave");
    constexpr int kStride = kInterleaveRows ? kWave / G : 1;
    constexpr int kWavesPerBlock = kBlock / kWave;
    int wave = (int)((blockIdx.x * kBlock + threadIdx.x) / kWave);
    if (reverse)
        wave = (int)gridDim.x * kWavesPerBlock - 1 - wave;
    const int gw = (int)(threadIdx.x % kWave) / G;
    const int row0 = kInterleaveRows ? wave * (kWave / G) * RPG + gw : (wave * (kWave / G) + gw) * RPG;
    const int lane = threadIdx.x % G;
    if (row0 >= nrow)
        return;
    RowAcc<S, BR * RPG> acc;
    if (alpha != S(0)) {
        acc = group_rows_product<S, BR, BC, G, RPG>(row0, nrow, lane, offsets, row_counts, columns, values, x,
                                                    kStride);
    } else {
#pragma unroll
        for (int s = 0; s < BR * RPG; ++s)
            acc.v[s] = S(0);
    }
    for (int e = lane; e < BR * RPG && row0 + (e / BR) * kStride < nrow; e += G) {
        const size_t yi = (size_t)(row0 + (e / BR) * kStride) * BR + e % BR;
        S v = pick(acc, e) * alpha;
        if (beta != S(0))
            v = fma(beta, y[yi], v);
        y[yi] = v;
    }
}

// Any block shape: one thread per scalar row, blocks in storage order (the
// reference row kernel's order).
template <typename S>
__global__ __launch_bounds__(kBlock) void bsr_mv_generic_kernel(
    int nrow, int br, int bc, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y,
    S alpha, S beta
)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)nrow * br)
        return;
    const int row = (int)(t / br);
    const int s = (int)(t % br);
    S v = S(0);
    if (alpha != S(0)) {
        const int end = row_end(offsets, row_counts, row);
        for (int b = offsets[row]; b < end; ++b) {
            const S* blk = values + (size_t)b * br * bc + (size_t)s * bc;
            const S* xs = x + (size_t)columns[b] * bc;
            for (int c = 0; c < bc; ++c)
                v = fma(blk[c], xs[c], v);
        }
        v *= alpha;
    }
    if (beta != S(0))
        v = fma(beta, y[t], v);
    y[t] = v;
}

template <typename S, int BR, int BC, int G, int RPG>
int launch_rows(
    int nrow, const int* offsets, const int* row_counts, const int* columns, const S* values, const S* x, S* y,
    S alpha, S beta, int reverse, hipStream_t stream
)
{
    constexpr int kRowsPerBlock = (kBlock / G) * RPG;
    const int64_t blocks = ceil_div(nrow, kRowsPerBlock);
    bsr_mv_rows_kernel<S, BR, BC, G, RPG><<<(unsigned)blocks, kBlock, 0, stream>>>(
        nrow, offsets, row_counts, columns, values, x, y, alpha, beta, reverse
    );
    WA_LAUNCH_CHECK();
    return kOk;
}

template <typename S, int BR, int BC, int G>
int launch_rpg(
    int rpg, int nrow, const int* offsets, const int* row_counts, const int* columns, const S* values, const S* x,
    S* y, S alpha, S beta, int reverse, hipStream_t stream
)
{
#define WA_RPG_ARGS nrow, offsets, row_counts, columns, values, x, y, alpha, beta, reverse, stream
    switch (rpg) {
    case 1: return launch_rows<S, BR, BC, G, 1>(WA_RPG_ARGS);
    case 2: return launch_rows<S, BR, BC, G, 2>(WA_RPG_ARGS);
    case 8:
        if constexpr (BR * BC == 1)
            return launch_rows<S, BR, BC, G, 8>(WA_RPG_ARGS);
        [[fallthrough]];
    default: return launch_rows<S, BR, BC, G, 4>(WA_RPG_ARGS);
    }
#undef WA_RPG_ARGS
}

template <typename S, int BR, int BC>
int launch_lanes(
    int g, int rpg, int nrow, const int* offsets, const int* row_counts, const int* columns, const S* values,
    const S* x, S* y, S alpha, S beta, int reverse, hipStream_t stream
)
{
#define WA_LANES_CASE(G)                                                                                              \
    case G:                                                                                                           \
        return launch_rpg<S, BR, BC, G>(rpg, nrow, offsets, row_counts, columns, values, x, y, alpha, beta, reverse, \
                                        stream);
    switch (g) {
        WA_LANES_CASE(1)
        WA_LANES_CASE(2)
        WA_LANES_CASE(4)
        WA_LANES_CASE(8)
        WA_LANES_CASE(16)
        WA_LANES_CASE(32)
        WA_LANES_CASE(64)
    default:
        set_error("bsr_mv: lanes per row must be a power of two in [1, 64], got %d", g);
        return kErrArg;
    }
#undef WA_LANES_CASE
}

template <typename S>
int bsr_mv_impl(
    int nrow, int ncol, int br, int bc, const int* offsets, const int* row_counts, const int* columns,
    const S* values, int64_t nnz_hint, const S* x, S* y, S alpha, S beta, int lanes, int rows_per_group,
    void* stream_ptr, int reverse = 0
)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && br > 0 && bc > 0, "bsr_mv: invalid shape (%d, %d, %d, %d)", nrow, ncol, br, bc);
    if (nrow == 0)
        return kOk;
    WA_REQUIRE(offsets && y, "bsr_mv: null offsets or y");
    WA_REQUIRE(alpha == S(0) || (columns && values && x), "bsr_mv: null matrix or x with nonzero alpha");
    WA_REQUIRE(rows_per_group <= 0 || rows_per_group == 1 || rows_per_group == 2 || rows_per_group == 4 ||
                   (rows_per_group == 8 && br * bc == 1),
               "bsr_mv: rows per group must be 1, 2 or 4 (or 8 for scalar blocks), got %d", rows_per_group);
    hipStream_t stream = as_stream(stream_ptr);
    const BsrShape shape = default_bsr_shape(nrow, br, bc, nnz_hint);
    const int g = lanes > 0 ? lanes : shape.lanes;
    const int rpg = rows_per_group > 0 ? rows_per_group : shape.rows_per_group;
    if (br == 1 && bc == 1)
        return launch_lanes<S, 1, 1>(g, rpg, nrow, offsets, row_counts, columns, values, x, y, alpha, beta, reverse,
                                     stream);
    if (br == 3 && bc == 3)
        return launch_lanes<S, 3, 3>(g, rpg, nrow, offsets, row_counts, columns, values, x, y, alpha, beta, reverse,
                                     stream);
    const int64_t threads = (int64_t)nrow * br;
    bsr_mv_generic_kernel<S><<<(unsigned)ceil_div(threads, kBlock), kBlock, 0, stream>>>(
        nrow, br, bc, offsets, row_counts, columns, values, x, y, alpha, beta
    );
    WA_LAUNCH_CHECK();
    return kOk;
}

// ---------------------------------------------------------------------------
// Diagonal blocks and Jacobi
// ---------------------------------------------------------------------------

// Index of the block (row, row) among the sorted active blocks of `row`, or -1
// (reference bsr_block_index, warp/_src/sparse.py:1881-1905).
__device__ __forceinline__ int diag_block_index(int row, const int* offsets, const int* row_counts, const int* columns)
{
    int lo = offsets[row];
    const int end = row_end(offsets, row_counts, row);
    int hi = end;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (columns[mid] < row)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < end && columns[lo] == row) ? lo : -1;
}

template <typename S>
__global__ __launch_bounds__(kBlock) void get_diag_kernel(
    int dim, int bsz, const int* offsets, const int* row_counts, const int* columns, const S* values, S scale, S* out
)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)dim * bsz)
        return;
    const int row = (int)(t / bsz);
    const int k = (int)(t % bsz);
    const int b = diag_block_index(row, offsets, row_counts, columns);
    out[t] = b >= 0 ? scale * values[(size_t)b * bsz + k] : S(0);
}

template <typename S> __device__ __forceinline__ S inverse_coefficient(S d, int use_abs)
{
    // 1/d (or 1/|d|), and 1 where d == 0 (linear.py:2283-2287)
    const S den = use_abs ? (d < S(0) ? -d : d) : d;
    return d == S(0) ? S(1) : S(1) / den;
}

// mode 0: inverse of the diagonal entries of each diagonal block (point
// Jacobi, out[nrow*br]).  mode 1: inverse of each br x br diagonal block
// (block Jacobi, out[nrow*br*br]); br == 3 uses the closed-form adjugate in
// FP64, and a singular or absent block falls back to the point inverse.
template <typename S>
__global__ __launch_bounds__(kBlock) void jacobi_kernel(
    int nrow, int dim, int br, const int* offsets, const int* row_counts, const int* columns, const S* values,
    int use_abs, int mode, S* out
)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= nrow)
        return;
    const int bsz = br * br;
    const int b = row < dim ? diag_block_index(row, offsets, row_counts, columns) : -1;
    const S* blk = b >= 0 ? values + (size_t)b * bsz : nullptr;
    if (mode == 0) {
        for (int s = 0; s < br; ++s)
            out[(size_t)row * br + s] = inverse_coefficient(blk ? blk[s * br + s] : S(0), use_abs);
        return;
    }
    S* o = out + (size_t)row * bsz;
    if (br == 3 && blk) {
        const double a00 = blk[0], a01 = blk[1], a02 = blk[2];
        const double a10 = blk[3], a11 = blk[4], a12 = blk[5];
        const double a20 = blk[6], a21 = blk[7], a22 = blk[8];
        const double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
        const double c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
        const double c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
        const double det = a00 * c00 + a01 * c10 + a02 * c20;
        if (det != 0.0 && isfinite(det)) {
            const double inv = 1.0 / det;
            o[0] = S(c00 * inv); o[1] = S(c01 * inv); o[2] = S(c02 * inv);
            o[3] = S(c10 * inv); o[4] = S(c11 * inv); o[5] = S(c12 * inv);
            o[6] = S(c20 * inv); o[7] = S(c21 * inv); o[8] = S(c22 * inv);
            return;
        }
    }
    if (br == 1 && blk) {
        o[0] = inverse_coefficient(blk[0], use_abs);
        return;
    }
    for (int s = 0; s < br; ++s)
        for (int c = 0; c < br; ++c)
            o[s * br + c] = s == c ? inverse_coefficient(blk ? blk[s * br + s] : S(0), use_abs) : S(0);
}

template <typename S>
int get_diag_impl(
    int nrow, int ncol, int br, int bc, const int* offsets, const int* row_counts, const int* columns,
    const S* values, S scale, S* out, void* stream
)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && br > 0 && bc > 0, "bsr_get_diag: invalid shape");
    const int dim = nrow < ncol ? nrow : ncol;
    if (dim == 0)
        return kOk;
    const int64_t threads = (int64_t)dim * br * bc;
    get_diag_kernel<S><<<(unsigned)ceil_div(threads, kBlock), kBlock, 0, as_stream(stream)>>>(
        dim, br * bc, offsets, row_counts, columns, values, scale, out
    );
    WA_LAUNCH_CHECK();
    return kOk;
}

template <typename S>
int jacobi_impl(
    int nrow, int ncol, int br, int bc, const int* offsets, const int* row_counts, const int* columns,
    const S* values, int use_abs, int mode, S* out, void* stream
)
{
    WA_REQUIRE(br == bc && br > 0, "bsr_jacobi: blocks must be square, got %dx%d", br, bc);
    WA_REQUIRE(mode == 0 || mode == 1, "bsr_jacobi: mode must be 0 (point) or 1 (block)");
    if (nrow == 0)
        return kOk;
    const int dim = nrow < ncol ? nrow : ncol;
    jacobi_kernel<S><<<(unsigned)ceil_div(nrow, kBlock), kBlock, 0, as_stream(stream)>>>(
        nrow, dim, br, offsets, row_counts, columns, values, use_abs, mode, out
    );
    WA_LAUNCH_CHECK();
    return kOk;
}

// ---------------------------------------------------------------------------
// Transposed product y = alpha A^T x + beta y (reference: bsr_mv with
// transpose=True, warp/_src/sparse.py:4718-4747, 4880-4895).  Like the
// reference, a scatter with atomic adds: the summation order of y's entries,
// and so their last bits, vary from run to run.
// ---------------------------------------------------------------------------

// y = beta * y (beta == 0: y = 0 without reading it)
template <typename S>
__global__ __launch_bounds__(kBlock) void scale_vector_kernel(int64_t n, S beta, S* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        y[i] = beta == S(0) ? S(0) : beta * y[i];
}

// G lanes per block row; each lane scatters alpha * A_b^T x_row into y.
template <typename S, int BR, int BC, int G>
__global__ __launch_bounds__(kBlock) void bsr_mv_transpose_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ row_counts, const int* __restrict__ columns,
    const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y, S alpha
)
{
    const int row = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    if (row >= nrow)
        return;
    const int beg = offsets[row];
    const int end = row_end(offsets, row_counts, row);
    S xr[BR];
#pragma unroll
    for (int r = 0; r < BR; ++r)
        xr[r] = alpha * x[(size_t)row * BR + r];
    for (int k = beg + lane; k < end; k += G) {
        const int col = columns[k];
        const S* blk = values + (size_t)k * (BR * BC);
#pragma unroll
        for (int c = 0; c < BC; ++c) {
            S sum = S(0);
#pragma unroll
            for (int r = 0; r < BR; ++r)
                sum = fma(blk[r * BC + c], xr[r], sum);
            atomicAdd(&y[(size_t)col * BC + c], sum);
        }
    }
}

// y = alpha A^T x + beta y without atomics, for a square compact matrix
// whose pattern is structurally symmetric (every block (r, c) has a mirror
// (c, r)): output block row c gathers, over row c's blocks (c, r) in storage
// order, the mirror block (r, c) -- found by binary search in row r -- times
// x[r].  G lanes per output row: lane g takes blocks g, g + G, ... with fused
// multiply-adds, then an xor butterfly; the result depends only on G.
template <typename S, int B, int G>
__global__ __launch_bounds__(kBlock) void bsr_mv_transpose_gather_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ columns, const S* __restrict__ values,
    const S* __restrict__ x, S* __restrict__ y, S alpha, S beta
)
{
    const int c = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    if (c >= nrow)
        return;
    const int beg = offsets[c];
    const int end = offsets[c + 1];
    S acc[B];
#pragma unroll
    for (int j = 0; j < B; ++j)
        acc[j] = S(0);
    for (int k = beg + lane; k < end; k += G) {
        const int r = columns[k];
        if (r < 0 || r >= nrow)
            continue;
        const int rend = offsets[r + 1];
        int lo = offsets[r], hi = rend;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (columns[mid] < c)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo == rend || columns[lo] != c)
            continue;  // no mirror block (the caller checks that there is always one)
        const S* blk = values + (size_t)lo * (B * B);  // block (r, c): B x B, row-major
#pragma unroll
        for (int j = 0; j < B; ++j)
#pragma unroll
            for (int i = 0; i < B; ++i)
                acc[j] = fma(blk[i * B + j], x[(size_t)r * B + i], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const S v = group_sum<G>(acc[j]);
        if (lane == 0) {
            S out = v * alpha;
            if (beta != S(0))
                out = fma(beta, y[(size_t)c * B + j], out);
            y[(size_t)c * B + j] = out;
        }
    }
}

template <typename S>
int bsr_mv_transpose_gather_impl(int nrow, int b, const int* offsets, const int* columns, const S* values,
                                 int64_t nnz_hint, const S* x, S* y, S alpha, S beta, hipStream_t stream)
{
    const int g = bsr_mv_lanes(nrow, b, b, nnz_hint);
#define WA_GATHER_CASE(B, G)                                                                                         \
    if (b == B && g == G) {                                                                                           \
        bsr_mv_transpose_gather_kernel<S, B, G><<<(unsigned)ceil_div(nrow, kBlock / G), kBlock, 0, stream>>>(        \
            nrow, offsets, columns, values, x, y, alpha, beta);                                                       \
        WA_LAUNCH_CHECK();                                                                                            \
        return kOk;                                                                                                   \
    }
    WA_GATHER_CASE(1, 1) WA_GATHER_CASE(1, 2) WA_GATHER_CASE(1, 4) WA_GATHER_CASE(1, 8)
    WA_GATHER_CASE(1, 16) WA_GATHER_CASE(1, 32) WA_GATHER_CASE(1, 64)
    WA_GATHER_CASE(3, 1) WA_GATHER_CASE(3, 2) WA_GATHER_CASE(3, 4) WA_GATHER_CASE(3, 8)
    WA_GATHER_CASE(3, 16) WA_GATHER_CASE(3, 32) WA_GATHER_CASE(3, 64)
#undef WA_GATHER_CASE
    set_error("bsr_mv_transpose_gather: block size must be 1 or 3, got %d", b);
    return kErrArg;
}

// Any block shape: one thread per (block row, stored block).
template <typename S>
__global__ __launch_bounds__(kBlock) void bsr_mv_transpose_generic_kernel(
    int nrow, int br, int bc, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y, S alpha
)
{
    const int row = blockIdx.x;
    if (row >= nrow)
        return;
    const int beg = offsets[row];
    const int end = row_end(offsets, row_counts, row);
    for (int k = beg + (int)threadIdx.x; k < end; k += kBlock) {
        const int col = columns[k];
        const S* blk = values + (size_t)k * br * bc;
        for (int c = 0; c < bc; ++c) {
            S sum = S(0);
            for (int r = 0; r < br; ++r)
                sum = fma(blk[r * bc + c], alpha * x[(size_t)row * br + r], sum);
            atomicAdd(&y[(size_t)col * bc + c], sum);
        }
    }
}

template <typename S>
int bsr_mv_transpose_impl(int nrow, int ncol, int br, int bc, const int* offsets, const int* row_counts,
                          const int* columns, const S* values, int64_t nnz_hint, const S* x, S* y, S alpha, S beta,
                          hipStream_t stream)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && br > 0 && bc > 0, "bsr_mv: bad shape");
    const int64_t ny = (int64_t)ncol * bc;
    if (ny > 0 && beta != S(1)) {
        scale_vector_kernel<S><<<(unsigned)ceil_div(ny, kBlock), kBlock, 0, stream>>>(ny, beta, y);
        WA_LAUNCH_CHECK();
    }
    if (alpha == S(0) || nrow == 0)
        return kOk;
    if (br == bc && (br == 1 || br == 3)) {
        const int g = bsr_mv_lanes(nrow, br, bc, nnz_hint);
#define WA_TRANSPOSE_CASE(B, G)                                                                                      \
    if (br == B && g == G) {                                                                                           \
        bsr_mv_transpose_kernel<S, B, B, G><<<(unsigned)ceil_div(nrow, kBlock / G), kBlock, 0, stream>>>(            \
            nrow, offsets, row_counts, columns, values, x, y, alpha                                                  \
        );                                                                                                           \
        WA_LAUNCH_CHECK();                                                                                           \
        return kOk;                                                                                                  \
    }
        WA_TRANSPOSE_CASE(1, 1) WA_TRANSPOSE_CASE(1, 2) WA_TRANSPOSE_CASE(1, 4) WA_TRANSPOSE_CASE(1, 8)
        WA_TRANSPOSE_CASE(1, 16) WA_TRANSPOSE_CASE(1, 32) WA_TRANSPOSE_CASE(1, 64)
        WA_TRANSPOSE_CASE(3, 1) WA_TRANSPOSE_CASE(3, 2) WA_TRANSPOSE_CASE(3, 4) WA_TRANSPOSE_CASE(3, 8)
        WA_TRANSPOSE_CASE(3, 16) WA_TRANSPOSE_CASE(3, 32) WA_TRANSPOSE_CASE(3, 64)
#undef WA_TRANSPOSE_CASE
    }
    bsr_mv_transpose_generic_kernel<S><<<(unsigned)nrow, kBlock, 0, stream>>>(nrow, br, bc, offsets, row_counts,
                                                                             columns, values, x, y, alpha);
    WA_LAUNCH_CHECK();
    return kOk;
}

}  // namespace

int bsr_mv_lanes(int nrow, int br, int bc, int64_t nnz_hint)
{
    if ((br == 1 && bc == 1) || (br == 3 && bc == 3))
        return default_bsr_shape(nrow, br, bc, nnz_hint).lanes;
    return 1;
}

}  // namespace wa

// Lanes per block row the bsr_mv / CG kernels use for this shape (1 for the
// one-thread-per-scalar-row kernel), so callers can reproduce the summation
// order: per lane, blocks g, g+G, ... in order, then an xor butterfly.
WA_API int warp_amd_bsr_mv_lanes(int nrow, int block_rows, int block_cols, int64_t nnz_hint)
{
    return wa::bsr_mv_lanes(nrow, block_rows, block_cols, nnz_hint);
}

#define WA_DEFINE_SPARSE_MV(SUFFIX, S)                                                                                \
    WA_API int warp_amd_bsr_mv_##SUFFIX(                                                                              \
        int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,                \
        const int* columns, const S* values, int64_t nnz_hint, const S* x, S* y, S alpha, S beta, void* stream        \
    )                                                                                                                 \
    {                                                                                                                 \
        return wa::bsr_mv_impl<S>(                                                                                    \
            nrow, ncol, block_rows, block_cols, offsets, row_counts, columns, values, nnz_hint, x, y, alpha, beta,    \
            0, 0, stream                                                                                              \
        );                                                                                                            \
    }                                                                                                                 \
    WA_API int warp_amd_bsr_mv_ex_##SUFFIX(                                                                           \
        int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,                \
        const int* columns, const S* values, int64_t nnz_hint, const S* x, S* y, S alpha, S beta, int lanes,          \
        int rows_per_group, void* stream                                                                              \
    )                                                                                                                 \
    {                                                                                                                 \
        return wa::bsr_mv_impl<S>(                                                                                    \
            nrow, ncol, block_rows, block_cols, offsets, row_counts, columns, values, nnz_hint, x, y, alpha, beta,    \
            lanes, rows_per_group, stream                                                                             \
        );                                                                                                            \
    }                                                                                                                 \
    WA_API int warp_amd_bsr_get_diag_##SUFFIX(                                                                        \
        int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,                \
        const int* columns, const S* values, S scale, S* out, void* stream                                            \
    )                                                                                                                 \
    {                                                                                                                 \
        return wa::get_diag_impl<S>(                                                                                  \
            nrow, ncol, block_rows, block_cols, offsets, row_counts, columns, values, scale, out, stream              \
        );                                                                                                            \
    }                                                                                                                 \
    WA_API int warp_amd_bsr_jacobi_##SUFFIX(                                                                          \
        int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,                \
        const int* columns, const S* values, int use_abs, int mode, S* out, void* stream                              \
    )                                                                                                                 \
    {                                                                                                                 \
        return wa::jacobi_impl<S>(                                                                                    \
            nrow, ncol, block_rows, block_cols, offsets, row_counts, columns, values, use_abs, mode, out, stream      \
        );                                                                                                            \
    }

WA_DEFINE_SPARSE_MV(f32, float)
WA_DEFINE_SPARSE_MV(f64, double)

// The matrix's arguments of a product, packed once by the caller (host
// struct; layout of warp_amd_bsr_mv_args in include/warp_amd.h).
struct WaBsrMvArgs {
    int nrow, ncol, block_rows, block_cols;
    const int* offsets;
    const int* row_counts;
    const int* columns;
    const void* values;
    int64_t nnz_hint;
    int lanes, rows_per_group, scalar_size, sweep;
};
static_assert(sizeof(WaBsrMvArgs) == 72, "warp_amd_bsr_mv_args layout");

WA_API int warp_amd_bsr_mv_packed(const WaBsrMvArgs* a, const void* x, void* y, double alpha, double beta,
                                  void* stream)
{
    WA_REQUIRE(a != nullptr, "bsr_mv_packed: null argument struct");
    if (a->scalar_size == 4)
        return wa::bsr_mv_impl<float>(a->nrow, a->ncol, a->block_rows, a->block_cols, a->offsets, a->row_counts,
                                      a->columns, static_cast<const float*>(a->values), a->nnz_hint,
                                      static_cast<const float*>(x), static_cast<float*>(y), (float)alpha,
                                      (float)beta, a->lanes, a->rows_per_group, stream, a->sweep & 1);
    WA_REQUIRE(a->scalar_size == 8, "bsr_mv_packed: scalar size must be 4 or 8, got %d", a->scalar_size);
    return wa::bsr_mv_impl<double>(a->nrow, a->ncol, a->block_rows, a->block_cols, a->offsets, a->row_counts,
                                   a->columns, static_cast<const double*>(a->values), a->nnz_hint,
                                   static_cast<const double*>(x), static_cast<double*>(y), alpha, beta, a->lanes,
                                   a->rows_per_group, stream, a->sweep & 1);
}

// y = alpha * A^T x + beta * y for an nrow x ncol BSR matrix of
// block_rows x block_cols blocks: x has nrow * block_rows entries, y has
// ncol * block_cols.  Atomic scatter, as the reference: non-deterministic
// last bits.
#define WA_DEFINE_MV_TRANSPOSE(SUFFIX, S)                                                                            \
    WA_API int warp_amd_bsr_mv_transpose_##SUFFIX(                                                                    \
        int nrow, int ncol, int block_rows, int block_cols, const int* offsets, const int* row_counts,                \
        const int* columns, const S* values, int64_t nnz_hint, const S* x, S* y, S alpha, S beta, void* stream        \
    )                                                                                                                 \
    {                                                                                                                 \
        return wa::bsr_mv_transpose_impl<S>(nrow, ncol, block_rows, block_cols, offsets, row_counts, columns,       \
                                            values, nnz_hint, x, y, alpha, beta, wa::as_stream(stream));             \
    }                                                                                                                 \
    WA_API int warp_amd_bsr_mv_transpose_gather_##SUFFIX(                                                             \
        int nrow, int block_size, const int* offsets, const int* columns, const S* values, int64_t nnz_hint,         \
        const S* x, S* y, S alpha, S beta, void* stream                                                               \
    )                                                                                                                 \
    {                                                                                                                 \
        WA_REQUIRE(nrow >= 0 && offsets && y, "bsr_mv_transpose_gather: invalid arguments");                         \
        if (nrow == 0)                                                                                                \
            return wa::kOk;                                                                                           \
        WA_REQUIRE(columns && values && x, "bsr_mv_transpose_gather: null arrays");                                  \
        return wa::bsr_mv_transpose_gather_impl<S>(nrow, block_size, offsets, columns, values, nnz_hint, x, y, alpha, \
                                                   beta, wa::as_stream(stream));                                      \
    }

WA_DEFINE_MV_TRANSPOSE(f32, float)
WA_DEFINE_MV_TRANSPOSE(f64, double)
