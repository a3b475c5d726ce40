// Device-side building blocks for block-sparse row (BSR) products.
//
// Storage is the compact / padded BSR layout of the reference
// (warp/_src/sparse.py:103-131): offsets[nrow+1], optional row_counts[nrow]
// (active blocks of row r are offsets[r] .. offsets[r]+row_counts[r]),
// columns[], and values[] holding BR x BC row-major scalars per block.
#pragma once

#include "common.h"

namespace wa {

__device__ __forceinline__ int row_end(const int* offsets, const int* row_counts, int row)
{
    return row_counts ? offsets[row] + row_counts[row] : offsets[row + 1];
}

// Per-lane partial sums of one block row (kept in registers: only ever
// indexed with compile-time indices).
template <typename S, int BR> struct RowAcc {
    S v[BR];
};

// One block row handled by an aligned group of G lanes (vector-BSR): lane g
// takes blocks beg+g, beg+g+G, ..., accumulates its BR partial row sums with
// explicit FMAs (block columns in order), and the group then reduces them with
// an xor butterfly (lane g adds lane g^off for off = G/2 .. 1).  On return every
// lane of the group holds the BR sums of (A x)[row*BR + s].  The order is
// fixed by G alone, so a CPU restatement can reproduce the bits.
template <typename S, int BR, int BC, int G>
__device__ __forceinline__ RowAcc<S, BR> group_row_product(
    int row, int lane, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const S* __restrict__ values, const S* __restrict__ x
)
{
    RowAcc<S, BR> acc;
#pragma unroll
    for (int s = 0; s < BR; ++s)
        acc.v[s] = S(0);
    const int beg = offsets[row];
    const int end = row_end(offsets, row_counts, row);
    for (int b = beg + lane; b < end; b += G) {
        const int col = columns[b];
        const S* blk = values + (size_t)b * (BR * BC);
        S m[BR * BC];
#pragma unroll
        for (int k = 0; k < BR * BC; ++k)
            m[k] = blk[k];
        S xv[BC];
#pragma unroll
        for (int c = 0; c < BC; ++c)
            xv[c] = x[(size_t)col * BC + c];
#pragma unroll
        for (int s = 0; s < BR; ++s)
#pragma unroll
            for (int c = 0; c < BC; ++c)
                acc.v[s] = fma(m[s * BC + c], xv[c], acc.v[s]);
    }
#pragma unroll
    for (int s = 0; s < BR; ++s)
        acc.v[s] = group_sum<G>(acc.v[s]);
    return acc;
}

// Several consecutive block rows per lane group: lane g takes block g, g+G, ...
// of each of the RPG rows row0 .. row0+RPG-1, issuing the loads of all rows
// before the FMAs (more bytes in flight per wave).  Each row is summed in the
// same order as group_row_product, so results are identical to it.  On return
// acc.v[j*BR + s] holds (A x)[(row0+j*rstride)*BR + s] in every lane of the
// group; rows at or beyond nrow contribute zeros.
//
// V is the stored value type and S the vector / accumulation type (V = float,
// S = double gives the FP64 residual of an FP32 matrix).
//
// XF: x(i) returns entry i of the vector being multiplied (a plain load, or
// a value computed from several arrays by fused kernels).
template <typename S, int BR, int BC, int G, int RPG, typename V, typename XF>
__device__ __forceinline__ RowAcc<S, BR * RPG> group_rows_product_f(
    int row0, int nrow, int lane, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const V* __restrict__ values, XF x, int rstride = 1
)
{
    RowAcc<S, BR * RPG> acc;
#pragma unroll
    for (int s = 0; s < BR * RPG; ++s)
        acc.v[s] = S(0);
    int beg[RPG], cnt[RPG];
    int nmax = 0;
#pragma unroll
    for (int j = 0; j < RPG; ++j) {
        const int r = row0 + j * rstride;
        beg[j] = 0;
        cnt[j] = 0;
        if (r < nrow) {
            beg[j] = offsets[r];
            cnt[j] = row_end(offsets, row_counts, r) - beg[j];
        }
        nmax = cnt[j] > nmax ? cnt[j] : nmax;
    }
    for (int k = lane; k < nmax; k += G) {
        bool ok[RPG];
        int col[RPG];
        V m[RPG][BR * BC];
#pragma unroll
        for (int j = 0; j < RPG; ++j) {
            ok[j] = k < cnt[j];
            // Inactive slots read storage slot 0 (in bounds whenever any row
            // has blocks) and gather x at column 0: slot 0 of a padded matrix
            // may be an unused slot whose column is -1.
            const int b = ok[j] ? beg[j] + k : 0;
            const int cb = columns[b];
            col[j] = ok[j] ? cb : 0;
            const V* blk = values + (size_t)b * (BR * BC);
#pragma unroll
            for (int q = 0; q < BR * BC; ++q)
                m[j][q] = blk[q];
        }
#pragma unroll
        for (int j = 0; j < RPG; ++j) {
            S xv[BC];
#pragma unroll
            for (int c = 0; c < BC; ++c)
                xv[c] = x((size_t)col[j] * BC + c);
#pragma unroll
            for (int s = 0; s < BR; ++s) {
                S a = acc.v[j * BR + s];
#pragma unroll
                for (int c = 0; c < BC; ++c)
                    a = fma((S)m[j][s * BC + c], xv[c], a);
                acc.v[j * BR + s] = ok[j] ? a : acc.v[j * BR + s];
            }
        }
    }
#pragma unroll
    for (int s = 0; s < BR * RPG; ++s)
        acc.v[s] = group_sum<G>(acc.v[s]);
    return acc;
}

template <typename S> struct LoadX {
    const S* __restrict__ p;
    __device__ __forceinline__ S operator()(size_t i) const { return p[i]; }
};

template <typename S, int BR, int BC, int G, int RPG, typename V = S>
__device__ __forceinline__ RowAcc<S, BR * RPG> group_rows_product(
    int row0, int nrow, int lane, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const V* __restrict__ values, const S* __restrict__ x, int rstride = 1
)
{
    return group_rows_product_f<S, BR, BC, G, RPG, V>(row0, nrow, lane, offsets, row_counts, columns, values,
                                                      LoadX<S>{x}, rstride);
}

// acc.v[i] for a lane-dependent i.  Written as an explicit select chain on
// named values: a loop of selects gets turned back into a dynamic index,
// which moves the accumulators to LDS.
template <typename S, int N> __device__ __forceinline__ S pick(const RowAcc<S, N>& acc, int i)
{
    static_assert(N >= 1 && N <= 16, "pick supports up to 16 values");
    S v = acc.v[0];
#define WA_PICK_STEP(K)                                                                                               \
    if constexpr (N > K) {                                                                                            \
        const S t = acc.v[K];                                                                                         \
        v = i == K ? t : v;                                                                                           \
    }
    WA_PICK_STEP(1) WA_PICK_STEP(2) WA_PICK_STEP(3) WA_PICK_STEP(4) WA_PICK_STEP(5) WA_PICK_STEP(6) WA_PICK_STEP(7)
    WA_PICK_STEP(8) WA_PICK_STEP(9) WA_PICK_STEP(10) WA_PICK_STEP(11) WA_PICK_STEP(12) WA_PICK_STEP(13)
    WA_PICK_STEP(14) WA_PICK_STEP(15)
#undef WA_PICK_STEP
    return v;
}

// Rows per lane group of the vector-BSR kernels (measured on MI355X: two
// 3x3 rows, four scalar rows).
constexpr int rows_per_group_for(int br, int bc) { return (br == 1 && bc == 1) ? 4 : 2; }

// Lanes per block row and rows per lane group for a matrix with `nnz_hint`
// blocks over `nrow` rows: about one block per lane for block rows, two
// entries per lane for scalar rows (lanes = the power of two covering that).
struct BsrShape {
    int lanes;
    int rows_per_group;
};

inline BsrShape default_bsr_shape(int nrow, int br, int bc, int64_t nnz_hint)
{
    const double mean = nrow > 0 && nnz_hint > 0 ? (double)nnz_hint / nrow : 8.0;
    const double per_lane = (br == 1 && bc == 1) ? 2.0 : 1.0;
    int g = 1;
    while (g * per_lane < mean && g < 64)
        g <<= 1;
    return BsrShape{g, rows_per_group_for(br, bc)};
}

}  // namespace wa
