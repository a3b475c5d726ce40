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

// acc.v[i] for a lane-dependent i.  Written as explicit selects on named
// values: a loop of selects gets turned back into a dynamic index, which moves
// the accumulators to LDS.
template <typename S, int N> __device__ __forceinline__ S pick(const RowAcc<S, N>& acc, int i)
{
    static_assert(N >= 1 && N <= 4, "pick supports up to 4 rows per block");
    if constexpr (N == 1) {
        return acc.v[0];
    } else if constexpr (N == 2) {
        const S a = acc.v[0], b = acc.v[1];
        return i == 1 ? b : a;
    } else if constexpr (N == 3) {
        const S a = acc.v[0], b = acc.v[1], c = acc.v[2];
        return i == 2 ? c : (i == 1 ? b : a);
    } else {
        const S a = acc.v[0], b = acc.v[1], c = acc.v[2], d = acc.v[3];
        return i == 3 ? d : (i == 2 ? c : (i == 1 ? b : a));
    }
}

}  // namespace wa
