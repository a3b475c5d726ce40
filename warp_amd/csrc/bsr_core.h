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

// One block row handled by an aligned group of G lanes (vector-BSR): lane g
// takes blocks beg+g, beg+g+G, ..., accumulates its BR partial row sums, and
// the group then reduces them with cross-lane shuffles.  On return every lane
// of the group holds the BR sums of (A x)[row*BR + s].
template <typename S, int BR, int BC, int G>
__device__ __forceinline__ void group_row_product(
    int row, int lane, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const S* __restrict__ values, const S* __restrict__ x, S (&acc)[BR]
)
{
#pragma unroll
    for (int s = 0; s < BR; ++s)
        acc[s] = S(0);
    const int beg = offsets[row];
    const int end = row_end(offsets, row_counts, row);
    for (int b = beg + lane; b < end; b += G) {
        const int col = columns[b];
        const S* blk = values + (size_t)b * (BR * BC);
        S xv[BC];
#pragma unroll
        for (int c = 0; c < BC; ++c)
            xv[c] = x[(size_t)col * BC + c];
#pragma unroll
        for (int s = 0; s < BR; ++s)
#pragma unroll
            for (int c = 0; c < BC; ++c)
                acc[s] += blk[s * BC + c] * xv[c];
    }
#pragma unroll
    for (int s = 0; s < BR; ++s)
        acc[s] = group_sum<G>(acc[s]);
}

// Select acc[i] for a lane-dependent i without dynamic register indexing.
template <typename S, int N> __device__ __forceinline__ S pick(const S (&acc)[N], int i)
{
    S v = acc[0];
#pragma unroll
    for (int s = 1; s < N; ++s)
        v = (i == s) ? acc[s] : v;
    return v;
}

}  // namespace wa
