// COO triplets -> compact BSR on gfx950.
//
// Replaces wp_bsr_matrix_from_triplets_device (warp/native/sparse.cu:1168-1322,
// declared at warp/native/warp.h:563-602) and the value accumulation kernel
// _bsr_accumulate_triplet_values (warp/_src/sparse.py:1073-1090, 1122-1141).
//
// Algorithm (an MSD radix sort whose digits are a coarse row bucket, the row,
// then (column, triplet index)):
//   1. classify, per tile of triplets: drop out-of-range / masked-out /
//      all-zero triplets and count the tile's triplets per coarse bucket of
//      rows in LDS;
//   2. exclusive scan of the (bucket, tile) counts;
//   3. scatter key = (col << 32) | triplet_index of every triplet into its
//      bucket's segment for the tile (LDS-atomic ranks, no global atomics);
//   4. per bucket: count its rows in LDS -> row starts, and move its keys into
//      row order;
//   5. sort every row by key (bitonic in registers across one wavefront for
//      <= 64 keys, in LDS by one workgroup beyond, with global rank merges past
//      4096 keys): exactly the order of a stable sort by (row, col), whatever
//      order the scatters produced; unique count per row -> scan -> offsets;
//   6. per row (<= 64 keys: sorted again in registers): columns, inclusive run
//      ends (summed_block_offsets) when asked for, and per unique block
//      values = v[i0] + v[i1] + ... sequentially in ascending triplet order
//      (bit-identical to the reference CUDA order).
// Matrices of more than 4096 * 4096 rows use a single-level row histogram and
// scatter with global atomics for steps 1-4.
#include "common.h"

#include <algorithm>
#include <climits>

namespace wa {
namespace {

constexpr int kBlock = 256;
constexpr int kMediumRegs = 8;                     // keys per lane of the longest medium rows
constexpr int kMediumCap = kMediumRegs * kWave;    // rows sorted in one wave's registers
constexpr int kLargeCap = 4096;                    // rows sorted in one workgroup's LDS
constexpr uint64_t kPadKey = ~uint64_t(0);

// Where each row's keys live in the key array: at start[r], count[r] of them
// (count NULL: up to start[r + 1]).
struct RowRanges {
    const int* start;
    const int* count;
    __device__ __forceinline__ int base(int r) const { return start[r]; }
    __device__ __forceinline__ int size(int r) const { return count ? count[r] : start[r + 1] - start[r]; }
};

// Two-level row bucketing: rows are grouped into coarse buckets of 1 << shift
// consecutive rows (at most kMaxBuckets buckets, at most kMaxBucketRows rows
// each); the triplets are cut into tiles of `tile` consecutive triplets.
constexpr int kMaxBuckets = 4096;
constexpr int kMaxBucketRows = 8192;
constexpr int kMaxTiles = 4096;
constexpr int kMinTile = 16384;
// With at most kStagedBuckets buckets, the scatter stages each sub-tile of
// kSubTile triplets in LDS ordered by bucket and writes every bucket's run
// contiguously (~kSubTile / buckets keys per run) instead of key by key.
constexpr int kStagedBuckets = 256;
constexpr int kSubTile = 4096;

struct Partition {
    int shift = 0;
    int nb = 1;       // coarse buckets
    int tile = kMinTile;
    int ntiles = 1;
    bool tiled = false;   // false: single-level bucketing with global atomics
    bool staged = false;  // LDS-staged, run-contiguous scatter (nb <= kStagedBuckets)
};

Partition partition_of(int nrow, int T)
{
    Partition p;
    const auto buckets = [&](int shift) { return ((int64_t)nrow + (1ll << shift) - 1) >> shift; };
    while (buckets(p.shift) > kStagedBuckets && (1 << (p.shift + 1)) <= kMaxBucketRows)
        ++p.shift;
    p.staged = buckets(p.shift) <= kStagedBuckets;
    while (buckets(p.shift) > kMaxBuckets)
        ++p.shift;
    p.nb = (int)buckets(p.shift);
    if (p.nb < 1)
        p.nb = 1;
    p.tiled = (1 << p.shift) <= kMaxBucketRows && T > 0 && nrow > 0;
    p.staged = p.staged && p.tiled;
    const int64_t per = (T + kMaxTiles - 1) / kMaxTiles;
    p.tile = (int)(per > kMinTile ? (per + kBlock - 1) / kBlock * kBlock : kMinTile);
    p.ntiles = (int)((T + (int64_t)p.tile - 1) / p.tile);
    if (p.ntiles < 1)
        p.ntiles = 1;
    return p;
}

// row-length list counters (two used) in the workspace
constexpr int kListCounters = 64;

struct TripletWs {
    int* trow;
    int* hist;
    int* row_start;
    uint64_t* keys;
    uint64_t* alt;
    int* brow;
    int* tile_offsets;
    int* large_rows;   // rows of more than kMediumCap keys (count in counters[0])
    int* medium_rows;  // rows of kWave + 1 .. kMediumCap keys (count in counters[1])
    int* counters;
    void* scan_ws;
    int* summed_offsets;
    int* summed_indices;
    Partition part;
};

TripletWs carve_ws(void* base, size_t bytes, int nrow, int T, bool own_summed, size_t* needed)
{
    Carve c(base, bytes);
    TripletWs w;
    w.part = partition_of(nrow, T);
    const int64_t ntile_hist = w.part.tiled ? (int64_t)w.part.nb * w.part.ntiles + 1 : 1;
    w.trow = c.take<int>((size_t)T);
    w.hist = c.take<int>((size_t)nrow + 1);
    w.row_start = c.take<int>((size_t)nrow + 1);
    w.keys = c.take<uint64_t>((size_t)T);
    w.alt = c.take<uint64_t>((size_t)T);
    w.brow = w.part.tiled ? c.take<int>((size_t)T) : nullptr;
    w.tile_offsets = c.take<int>((size_t)ntile_hist);
    w.large_rows = c.take<int>((size_t)nrow + 1);
    w.medium_rows = c.take<int>((size_t)nrow + 1);
    w.counters = c.take<int>(kListCounters);
    w.scan_ws = c.take<char>(scan_workspace_bytes(std::max<int64_t>((int64_t)nrow + 1, ntile_hist)));
    w.summed_offsets = own_summed ? c.take<int>((size_t)T) : nullptr;
    w.summed_indices = own_summed ? c.take<int>((size_t)T) : nullptr;
    *needed = c.used;
    return w;
}

__global__ void fill_i32_kernel(int* p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Scalar pattern test of one block: true when every scalar has no bit in `mask`
// (sparse.cu:245-268; sparse.cpp:22-29).
template <typename U> __device__ __forceinline__ bool block_is_zero(const U* v, int block_size, U mask)
{
    for (int k = 0; k < block_size; ++k)
        if ((v[k] & mask) != U(0))
            return false;
    return true;
}

// Existing-topology test for masked assembly (sparse.cu:277-310).
__device__ __forceinline__ bool in_mask(int row, int col, const int* offsets, const int* row_counts, const int* columns)
{
    int lo = offsets[row];
    int hi = row_counts ? lo + row_counts[row] : offsets[row + 1];
    const int end = hi;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (columns[mid] < col)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < end && columns[lo] == col;
}

template <typename U>
__global__ __launch_bounds__(kBlock) void classify_kernel(
    int T, const int* tpl_count, int nrow, int ncol, const int* rows, const int* cols, const U* values,
    int block_size, U zero_mask, const int* mask_offsets, const int* mask_row_counts, const int* mask_columns,
    int* trow, int* hist
)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T)
        return;
    const int limit = tpl_count ? *tpl_count : T;
    const int r = rows[i];
    const int c = cols[i];
    bool valid = i < limit && r >= 0 && r < nrow && c >= 0 && c < ncol;
    if (valid && mask_offsets)
        valid = in_mask(r, c, mask_offsets, mask_row_counts, mask_columns);
    if (valid && values)
        valid = !block_is_zero(values + (size_t)i * block_size, block_size, zero_mask);
    trow[i] = valid ? r : -1;
    if (valid)
        atomicAdd(&hist[r], 1);
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(
    int T, const int* trow, const int* cols, const int* row_start, int* hist, uint64_t* keys
)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T)
        return;
    const int r = trow[i];
    if (r < 0)
        return;
    const int pos = row_start[r] + atomicSub(&hist[r], 1) - 1;
    keys[pos] = ((uint64_t)(uint32_t)cols[i] << 32) | (uint32_t)i;
}

// Tiled classification (one workgroup per tile): validity as classify_kernel,
// trow[t] = row or -1, and the tile's histogram of coarse buckets counted in
// LDS, stored bucket-major: tile_hist[b * ntiles + tile].
template <typename U>
__global__ __launch_bounds__(kBlock) void tile_classify_kernel(
    int T, const int* tpl_count, int nrow, int ncol, const int* rows, const int* cols, const U* values,
    int block_size, U zero_mask, const int* mask_offsets, const int* mask_row_counts, const int* mask_columns,
    Partition part, int* trow, int* tile_hist
)
{
    __shared__ int hist[kMaxBuckets];
    for (int b = threadIdx.x; b < part.nb; b += kBlock)
        hist[b] = 0;
    __syncthreads();
    const int limit = tpl_count ? *tpl_count : T;
    const int64_t t0 = (int64_t)blockIdx.x * part.tile;
    const int64_t t1 = t0 + part.tile < T ? t0 + part.tile : T;
    // kClassifyBatch triplets per thread per step, their rows, columns and the
    // first scalar of each block loaded together (one memory round trip); the
    // rest of a block is read only when its first scalar is zero
    constexpr int kClassifyBatch = 4;
    for (int64_t tb = t0 + threadIdx.x; tb < t1; tb += kClassifyBatch * kBlock) {
        int r[kClassifyBatch], c[kClassifyBatch];
        U v0[kClassifyBatch];
#pragma unroll
        for (int q = 0; q < kClassifyBatch; ++q) {
            const int64_t t = tb + (int64_t)q * kBlock;
            const bool in = t < t1;
            r[q] = in ? rows[t] : -1;
            c[q] = in ? cols[t] : -1;
            v0[q] = (in && values) ? values[(size_t)t * block_size] : U(0);
        }
#pragma unroll
        for (int q = 0; q < kClassifyBatch; ++q) {
            const int64_t t = tb + (int64_t)q * kBlock;
            if (t >= t1)
                break;
            bool valid = t < limit && r[q] >= 0 && r[q] < nrow && c[q] >= 0 && c[q] < ncol;
            if (valid && mask_offsets)
                valid = in_mask(r[q], c[q], mask_offsets, mask_row_counts, mask_columns);
            if (valid && values && (v0[q] & zero_mask) == U(0))
                valid = !block_is_zero(values + (size_t)t * block_size, block_size, zero_mask);
            trow[t] = valid ? r[q] : -1;
            if (valid)
                atomicAdd(&hist[r[q] >> part.shift], 1);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < part.nb; b += kBlock)
        tile_hist[(size_t)b * part.ntiles + blockIdx.x] = hist[b];
}

// Tiled scatter into coarse buckets (one workgroup per tile): each valid
// triplet's key (col << 32 | t) and row go to its bucket's segment of the
// tile, ranks taken with LDS atomics from the scanned tile histograms.
__global__ __launch_bounds__(kBlock) void tile_scatter_kernel(
    int T, const int* __restrict__ trow, const int* __restrict__ cols, Partition part,
    const int* __restrict__ tile_offsets, uint64_t* __restrict__ bkeys, int* __restrict__ brow,
    int* __restrict__ zero, int64_t nzero
)
{
    // zero[0 .. nzero): the row histogram of the next pass (no fill launch)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nzero; i += (int64_t)gridDim.x * kBlock)
        zero[i] = 0;
    __shared__ int cursor[kMaxBuckets];
    for (int b = threadIdx.x; b < part.nb; b += kBlock)
        cursor[b] = tile_offsets[(size_t)b * part.ntiles + blockIdx.x];
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * part.tile;
    const int64_t t1 = t0 + part.tile < T ? t0 + part.tile : T;
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBlock) {
        const int r = trow[t];
        if (r < 0)
            continue;
        const int pos = atomicAdd(&cursor[r >> part.shift], 1);
        bkeys[pos] = ((uint64_t)(uint32_t)cols[t] << 32) | (uint32_t)t;
        brow[pos] = r;
    }
}

// Tiled scatter with LDS staging (one workgroup per tile, nb <= 256 buckets):
// each sub-tile of kSubTile triplets is ranked per bucket in LDS, placed in
// LDS in bucket order, then copied out so that consecutive threads write
// consecutive keys of one bucket's run.  1024 threads per workgroup, four
// triplets each per sub-tile: the sub-tile's chain of LDS phases is bound by
// latency, which more waves per sub-tile hide (16.8 M shuffled element-loop
// triplets of a 64^3 hex mesh: 194 us with 256 threads, 104 with 1024).
constexpr int kStagedThreads = 1024;
__global__ __launch_bounds__(kStagedThreads) void tile_scatter_staged_kernel(
    int T, const int* __restrict__ trow, const int* __restrict__ cols, Partition part,
    const int* __restrict__ tile_offsets, uint64_t* __restrict__ bkeys, int* __restrict__ brow,
    int* __restrict__ zero, int64_t nzero
)
{
    // zero[0 .. nzero): the row histogram of the next pass (no fill launch)
    for (int64_t i = (int64_t)blockIdx.x * kStagedThreads + threadIdx.x; i < nzero;
         i += (int64_t)gridDim.x * kStagedThreads)
        zero[i] = 0;
    constexpr int PER = kSubTile / kStagedThreads;
    __shared__ int cursor[kStagedBuckets];
    __shared__ int lcnt[kStagedBuckets];
    __shared__ int loff[kStagedBuckets];
    __shared__ uint64_t skey[kSubTile];
    __shared__ int srow[kSubTile];
    const int tid = threadIdx.x;
    for (int b = tid; b < part.nb; b += kStagedThreads)
        cursor[b] = tile_offsets[(size_t)b * part.ntiles + blockIdx.x];
    const int64_t t0 = (int64_t)blockIdx.x * part.tile;
    const int64_t t1 = t0 + part.tile < T ? t0 + part.tile : T;
    for (int64_t sub = t0; sub < t1; sub += kSubTile) {
        for (int b = tid; b < kStagedBuckets; b += kStagedThreads)
            lcnt[b] = 0;
        __syncthreads();
        int r[PER], c[PER], rank[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int64_t t = sub + k * kStagedThreads + tid;
            r[k] = t < t1 ? trow[t] : -1;
            c[k] = r[k] >= 0 ? cols[t] : 0;
            rank[k] = r[k] >= 0 ? atomicAdd(&lcnt[r[k] >> part.shift], 1) : 0;
        }
        __syncthreads();
        if (tid < kWave) {
            // exclusive scan of the 256 bucket counts by one wavefront (4 per lane)
            int v[kStagedBuckets / kWave];
            int sum = 0;
#pragma unroll
            for (int j = 0; j < kStagedBuckets / kWave; ++j) {
                v[j] = lcnt[tid * (kStagedBuckets / kWave) + j];
                sum += v[j];
            }
            int incl = sum;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if (tid >= o)
                    incl += y;
            }
            int run = incl - sum;
#pragma unroll
            for (int j = 0; j < kStagedBuckets / kWave; ++j) {
                loff[tid * (kStagedBuckets / kWave) + j] = run;
                run += v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (r[k] < 0)
                continue;
            const int pos = loff[r[k] >> part.shift] + rank[k];
            const int64_t t = sub + k * kStagedThreads + tid;
            skey[pos] = ((uint64_t)(uint32_t)c[k] << 32) | (uint32_t)t;
            srow[pos] = r[k];
        }
        __syncthreads();
        const int n = loff[kStagedBuckets - 1] + lcnt[kStagedBuckets - 1];
        for (int i = tid; i < n; i += kStagedThreads) {
            const int row = srow[i];
            const int b = row >> part.shift;
            const int g = cursor[b] + (i - loff[b]);
            bkeys[g] = skey[i];
            brow[g] = row;
        }
        __syncthreads();
        for (int b = tid; b < part.nb; b += kStagedThreads)
            cursor[b] += lcnt[b];
    }
}

// Row pass over the bucket-ordered keys, one workgroup of kChunkThreads per
// chunk of kChunk keys: a chunk's rows lie in a window of kRowWindow rows from
// its first bucket, counted in LDS and added to the global histogram with one
// atomic per (chunk, row) (rows beyond the window: one atomic per key).  Each
// thread loads its kChunkPer keys up front and issues its atomics in unrolled
// batches: these kernels are bound by latency (a few thousand workgroups, one
// global round trip each), not by their ~270 MB of traffic.
constexpr int kChunk = 16384;
constexpr int kChunkThreads = 1024;
constexpr int kChunkPer = kChunk / kChunkThreads;
constexpr int kRowWindow = 8192;
constexpr int kWindowPer = kRowWindow / kChunkThreads;
constexpr int kXcds = 8;

// Workgroup b is dispatched to XCD b % kXcds (round-robin over the MI355X's 8
// XCDs); numbering the chunks so that each XCD takes a contiguous range keeps
// the chunks of one bucket -- which write into the same key range -- in one
// L2.  The grid is a multiple of kXcds; surplus workgroups find no chunk.
__device__ __forceinline__ int xcd_chunk()
{
    return (int)((blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds);
}

__global__ __launch_bounds__(kChunkThreads) void chunk_row_hist_kernel(int nrow, Partition part,
                                                                      const int* __restrict__ total,
                                                                      const int* __restrict__ brow,
                                                                      int* __restrict__ hist)
{
    __shared__ int cnt[kRowWindow];
    const int n = *total;
    const int i0 = xcd_chunk() * kChunk;
    if (i0 >= n)
        return;
    const int i1 = min(i0 + kChunk, n);
    const int rlo = (brow[i0] >> part.shift) << part.shift;
    int r[kChunkPer];
#pragma unroll
    for (int u = 0; u < kChunkPer; ++u) {
        const int i = i0 + (int)threadIdx.x + u * kChunkThreads;
        r[u] = i < i1 ? brow[i] : -1;
    }
#pragma unroll
    for (int q = 0; q < kWindowPer; ++q)
        cnt[threadIdx.x + q * kChunkThreads] = 0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kChunkPer; ++u) {
        if (r[u] < 0)
            continue;
        const int d = r[u] - rlo;
        if (d < kRowWindow)
            atomicAdd(&cnt[d], 1);
        else
            atomicAdd(&hist[r[u]], 1);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kWindowPer; ++q) {
        const int d = threadIdx.x + q * kChunkThreads;
        const int c = cnt[d];
        if (c && rlo + d < nrow)
            atomicAdd(&hist[rlo + d], c);
    }
}

// Scatter of the bucket-ordered keys into row order, one workgroup per chunk
// of kSChunk keys: the chunk's rows are counted in LDS (a window of
// kSWindow rows from its first bucket), each (chunk, row) run is reserved
// with one global atomic on fill[row] (zeroed beforehand), and the keys are
// placed in LDS in destination order -- runs in row order, keys within a run
// by LDS-atomic rank -- then written out by consecutive threads, so a run
// leaves as one contiguous store instead of one scattered store per key
// (rows beyond the window: one atomic per key).  On 16.8 M keys in 7.2 M
// blocks, keys scattered one by one from registers wrote 672 MB for 134 MB of
// keys (283 us); staged, 211 MB (114 us).
constexpr int kSChunk = 8192;
constexpr int kSPer = kSChunk / kChunkThreads;
constexpr int kSWindow = 4096;
constexpr int kSWinPer = kSWindow / kChunkThreads;

__global__ __launch_bounds__(kChunkThreads) void chunk_row_scatter_kernel(int nrow, Partition part,
                                                                         const int* __restrict__ total,
                                                                         const uint64_t* __restrict__ bkeys,
                                                                         const int* __restrict__ brow,
                                                                         const int* __restrict__ row_start,
                                                                         int* __restrict__ fill,
                                                                         uint64_t* __restrict__ keys,
                                                                         int* __restrict__ counters)
{
    // the row-length list counters of the next pass start at zero (no fill launch)
    if (blockIdx.x == 0 && threadIdx.x < kListCounters)
        counters[threadIdx.x] = 0;
    __shared__ int gbase[kSWindow];  // global start of the chunk's run of each window row
    __shared__ int lrank[kSWindow];  // keys per window row, then the running rank
    __shared__ int loff[kSWindow];   // start of each window row's run in the staged keys
    __shared__ uint64_t skey[kSChunk];
    __shared__ int spos[kSChunk];
    __shared__ int wsum[kChunkThreads / kWave];
    const int n = *total;
    const int i0 = xcd_chunk() * kSChunk;
    if (i0 >= n)
        return;
    const int i1 = min(i0 + kSChunk, n);
    const int rlo = (brow[i0] >> part.shift) << part.shift;
    int r[kSPer];
    uint64_t k[kSPer];
#pragma unroll
    for (int u = 0; u < kSPer; ++u) {
        const int i = i0 + (int)threadIdx.x + u * kChunkThreads;
        r[u] = i < i1 ? brow[i] : -1;
        k[u] = i < i1 ? bkeys[i] : 0;
    }
#pragma unroll
    for (int q = 0; q < kSWinPer; ++q)
        lrank[threadIdx.x + q * kChunkThreads] = 0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSPer; ++u) {
        const int d = r[u] - rlo;
        if (r[u] >= 0 && d < kSWindow)
            atomicAdd(&lrank[d], 1);
    }
    __syncthreads();
    // thread t owns window rows t * kSWinPer ..: their counts, global runs and
    // the exclusive scan of the counts over the window (thread order = row order)
    int c[kSWinPer];
    int tsum = 0;
#pragma unroll
    for (int q = 0; q < kSWinPer; ++q) {
        const int d = (int)threadIdx.x * kSWinPer + q;
        c[q] = rlo + d < nrow ? lrank[d] : 0;
        tsum += c[q];
    }
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
    int incl = tsum;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o)
            incl += y;
    }
    if (lane == kWave - 1)
        wsum[wv] = incl;
    __syncthreads();
    int run = incl - tsum;
    for (int w = 0; w < wv; ++w)
        run += wsum[w];
#pragma unroll
    for (int q = 0; q < kSWinPer; ++q) {
        const int d = (int)threadIdx.x * kSWinPer + q;
        loff[d] = run;
        run += c[q];
        gbase[d] = c[q] ? row_start[rlo + d] + atomicAdd(&fill[rlo + d], c[q]) : 0;
        lrank[d] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSPer; ++u) {
        if (r[u] < 0)
            continue;
        const int d = r[u] - rlo;
        if (d < kSWindow) {
            const int rk = atomicAdd(&lrank[d], 1);
            skey[loff[d] + rk] = k[u];
            spos[loff[d] + rk] = gbase[d] + rk;
        } else {
            keys[row_start[r[u]] + atomicAdd(&fill[r[u]], 1)] = k[u];
        }
    }
    __syncthreads();
    int staged = 0;
    for (int w = 0; w < kChunkThreads / kWave; ++w)
        staged += wsum[w];
    for (int i = threadIdx.x; i < staged; i += kChunkThreads)
        keys[spos[i]] = skey[i];
}

// In-LDS bitonic sort of `n` (power of two) keys by `nthreads` threads.
__device__ __forceinline__ void lds_bitonic(uint64_t* s, int n, int tid, int nthreads)
{
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < n / 2; t += nthreads) {
                const int lo = 2 * t - (t & (j - 1));
                const int hi = lo + j;
                const bool up = (lo & k) == 0;
                const uint64_t a = s[lo];
                const uint64_t b = s[hi];
                if ((a > b) == up) {
                    s[lo] = b;
                    s[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ int pow2_at_least(int n, int lo)
{
    int p = lo;
    while (p < n)
        p <<= 1;
    return p;
}

// Writes the sorted bucket back, the sorted triplet indices, and the unique
// column count of the row.
__device__ void emit_sorted_row(
    const uint64_t* s, int size, int base, int row, int tid, int nthreads, uint64_t* keys, int* summed_indices,
    int* urow, int* lds_count
)
{
    int heads = 0;
    for (int k = tid; k < size; k += nthreads) {
        const uint64_t key = s[k];
        keys[base + k] = key;
        summed_indices[base + k] = (int)(uint32_t)key;
        heads += (k == 0 || (s[k - 1] >> 32) != (key >> 32)) ? 1 : 0;
    }
    atomicAdd(lds_count, heads);
    __syncthreads();
    if (tid == 0)
        urow[row] = *lds_count;
}

// 64-bit key exchange with lane ^ mask.
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask)
{
    const int lo = __shfl_xor((int)(uint32_t)v, mask);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), mask);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Bitonic sort of one key per lane across the 64 lanes of a wavefront, in
// registers (no LDS): ascending by lane.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t key, int lane)
{
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t other = shfl_xor_u64(key, j);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            key = (lower == up) ? (key < other ? key : other) : (key < other ? other : key);
        }
    return key;
}

// Bitonic sort of one 32-bit value per lane across the wavefront (ascending by lane).
__device__ __forceinline__ unsigned wave_sort32(unsigned v, int lane)
{
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const unsigned other = (unsigned)__shfl_xor((int)v, j);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            v = (lower == up) ? (v < other ? v : other) : (v < other ? other : v);
        }
    return v;
}

// Rows of at most one key per lane: a row's keys, sorted in registers.
__device__ __forceinline__ uint64_t load_sorted_row(const uint64_t* keys, int base, int size, int lane)
{
    const uint64_t key = lane < size ? keys[base + lane] : kPadKey;
    return wave_sort64(key, lane);
}

// Rows are handled one per wavefront, kRowsPerBlock wavefronts per workgroup.
constexpr int kRowsPerBlock = 4;

// Lists the rows of more than kWave keys: those of at most kMediumCap keys in
// medium_rows (count in counters[1]), larger ones in large_rows (count in
// counters[0]).  One atomic per workgroup and list: a per-row append to one
// counter serialises ~10^5 device-scope atomics on a single address.
__global__ __launch_bounds__(kBlock) void classify_rows_kernel(
    int nrow, RowRanges rr, int* __restrict__ large_rows, int* __restrict__ medium_rows, int* __restrict__ counters
)
{
    __shared__ int s_cnt[2][kBlock / kWave];
    __shared__ int s_base[2];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int size = row < nrow ? rr.size(row) : 0;
    const bool medium = size > kWave && size <= kMediumCap;
    const bool large = size > kMediumCap;
    const uint64_t mm = __ballot(medium);
    const uint64_t ml = __ballot(large);
    if (lane == 0) {
        s_cnt[0][wv] = __popcll(mm);
        s_cnt[1][wv] = __popcll(ml);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        int total = 0;
        for (int w = 0; w < kBlock / kWave; ++w)
            total += s_cnt[threadIdx.x][w];
        // list 0 (medium) counts in counters[1], list 1 (large) in counters[0]
        s_base[threadIdx.x] = total ? atomicAdd(&counters[threadIdx.x == 0 ? 1 : 0], total) : 0;
    }
    __syncthreads();
    int off_m = s_base[0];
    int off_l = s_base[1];
    for (int w = 0; w < wv; ++w) {
        off_m += s_cnt[0][w];
        off_l += s_cnt[1][w];
    }
    const uint64_t below = (uint64_t(1) << lane) - 1;
    if (medium)
        medium_rows[off_m + __popcll(mm & below)] = row;
    if (large)
        large_rows[off_l + __popcll(ml & below)] = row;
}

// Unique column count of every row of at most kWave keys (sorted in
// registers; with write_back the sorted keys and triplet indices are stored
// in place).  Larger rows are listed by classify_rows_kernel.
__global__ __launch_bounds__(kWave * kRowsPerBlock) void count_small_rows_kernel(
    int nrow, RowRanges rr, uint64_t* __restrict__ keys, int* __restrict__ summed_indices,
    int write_back, int* __restrict__ urow
)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave);
    if (row >= nrow)
        return;
    const int base = rr.base(row);
    const int size = rr.size(row);
    if (size > kWave)
        return;
    if (write_back) {
        const uint64_t key = load_sorted_row(keys, base, size, lane);
        const int col = (int)(key >> 32);
        const int prev = __shfl_up(col, 1);
        const bool head = lane < size && (lane == 0 || prev != col);
        const int n = __popcll(__ballot(head));
        if (lane == 0)
            urow[row] = n;
        if (lane < size) {
            keys[base + lane] = key;
            summed_indices[base + lane] = (int)(uint32_t)key;
        }
        return;
    }
    // the count needs the columns only: sort their 32 bits (half the exchanges
    // of the 64-bit keys), past the row's end an all-ones column above every
    // real one
    const unsigned* khi = reinterpret_cast<const unsigned*>(keys);
    const unsigned col = wave_sort32(lane < size ? khi[2 * (size_t)(base + lane) + 1] : 0xffffffffu, lane);
    const unsigned prev = (unsigned)__shfl_up((int)col, 1);
    const bool head = lane < size && (lane == 0 || prev != col);
    const int n = __popcll(__ballot(head));
    if (lane == 0)
        urow[row] = n;
}

// Medium rows (kWave + 1 .. kMediumCap keys) live in R = 2, 4 or 8 registers
// per lane of one wave (the smallest R with R * kWave >= the row's keys).

// Bitonic sort of a lane-major register tile, element i = lane * R + q
// (kPadKey past the row's end), ascending by element index.  Every trip count
// is a compile-time constant, so the network unrolls with static register
// indices: exchanges at distance < R stay inside a lane, longer ones go across
// lanes (distance / R apart), which leaves fewer cross-lane steps than a
// register-major tile.
template <int R>
__device__ __forceinline__ void wave_sort_lane_major(uint64_t (&k)[R], int lane)
{
#pragma unroll
    for (int size = 2; size <= R * kWave; size <<= 1) {
#pragma unroll
        for (int j = size >> 1; j > 0; j >>= 1) {
            if (j < R) {
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    if (q & j)
                        continue;
                    const bool up = ((lane * R + q) & size) == 0;
                    const uint64_t a = k[q];
                    const uint64_t b = k[q | j];
                    const bool keep = (a < b) == up;
                    k[q] = keep ? a : b;
                    k[q | j] = keep ? b : a;
                }
            } else {
                const int lj = j / R;
                const bool lower = (lane & lj) == 0;
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const uint64_t other = shfl_xor_u64(k[q], lj);
                    const bool up = ((lane * R + q) & size) == 0;
                    // the lower element of an ascending pair keeps the smaller key
                    const bool take_other = (other < k[q]) == (lower == up);
                    k[q] = take_other ? other : k[q];
                }
            }
        }
    }
}

template <int R>
__device__ __forceinline__ void load_medium_row(const uint64_t* keys, int base, int size, int lane, uint64_t (&k)[R])
{
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = q * kWave + lane;
        k[q] = i < size ? keys[base + i] : kPadKey;
    }
}

// Column of element i - 1 (q, lane) of a register tile (-1 before the first).
template <int R>
__device__ __forceinline__ int prev_col(const int (&c)[R], int q, int lane)
{
    const int up = __shfl_up(c[q], 1);
    const int wrap = __shfl(c[q > 0 ? q - 1 : 0], kWave - 1);
    return lane > 0 ? up : (q > 0 ? wrap : -1);
}

// Column of element i + 1 (q, lane) of a register tile (-1 after the last).
template <int R>
__device__ __forceinline__ int next_col(const int (&c)[R], int q, int lane)
{
    const int down = __shfl_down(c[q], 1);
    const int wrap = __shfl(c[q < R - 1 ? q + 1 : q], 0);
    return lane < kWave - 1 ? down : (q < R - 1 ? wrap : -1);
}

// Orders the LDS writes of a wave before its own later LDS reads (no
// workgroup barrier: the waves of a workgroup work on different rows).
__device__ __forceinline__ void wave_lds_sync()
{
    // a wave's LDS instructions execute in issue order: only the compiler's
    // reordering has to be fenced (and no global loads in flight waited for)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One medium row (at most R * kWave keys): loaded lane-major, sorted in
// registers, stored back in place with its sorted triplet indices; returns its
// unique column count.
template <int R>
__device__ __forceinline__ int sort_medium_row(uint64_t* keys, int* summed_indices, int base, int size, int lane)
{
    uint64_t k[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = lane * R + q;
        k[q] = i < size ? keys[base + i] : kPadKey;
    }
    wave_sort_lane_major<R>(k, lane);
    const int prev_lane_last = __shfl_up((int)(k[R - 1] >> 32), 1);
    int nu = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = lane * R + q;
        const int c = (int)(k[q] >> 32);
        const int prev = q > 0 ? (int)(k[q - 1] >> 32) : prev_lane_last;
        const bool head = i < size && (i == 0 || prev != c);
        nu += __popcll(__ballot(head));
        if (i < size) {
            keys[base + i] = k[q];
            summed_indices[base + i] = (int)(uint32_t)k[q];
        }
    }
    return nu;
}

// Medium rows, one wave per listed row (grid-stride over the list).
__global__ __launch_bounds__(kWave * kRowsPerBlock) void sort_medium_rows_kernel(
    RowRanges rr, uint64_t* __restrict__ keys, int* __restrict__ summed_indices, int* __restrict__ urow,
    const int* __restrict__ medium_rows, const int* __restrict__ counters
)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int nmed = counters[1];
    const int nwaves = gridDim.x * kRowsPerBlock;
    for (int li = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave); li < nmed; li += nwaves) {
        const int row = medium_rows[li];
        const int base = rr.base(row);
        const int size = rr.size(row);
        const int nu = size <= 2 * kWave   ? sort_medium_row<2>(keys, summed_indices, base, size, lane)
                       : size <= 4 * kWave ? sort_medium_row<4>(keys, summed_indices, base, size, lane)
                                           : sort_medium_row<kMediumRegs>(keys, summed_indices, base, size, lane);
        if (lane == 0)
            urow[row] = nu;
    }
}

// Topology of one sorted medium row: columns and (optionally) inclusive run
// ends to global memory, run starts and sorted triplet indices to the wave's
// LDS.  Returns the number of unique columns.
template <int R>
__device__ __forceinline__ int medium_row_runs(const uint64_t* keys, int base, int size, int out0, int lane,
                                               int* bsr_columns, int* summed_offsets, int* idx_of, int* beg_of)
{
    const uint64_t below = lane == 63 ? ~uint64_t(0) : ((uint64_t(1) << (lane + 1)) - 1);
    uint64_t k[R];
    load_medium_row<R>(keys, base, size, lane, k);
    int c[R];
#pragma unroll
    for (int q = 0; q < R; ++q)
        c[q] = (int)(k[q] >> 32);
    int runs = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = q * kWave + lane;
        const bool act = i < size;
        const int prev = prev_col<R>(c, q, lane);
        const int next = next_col<R>(c, q, lane);
        const bool head = act && (i == 0 || prev != c[q]);
        const bool tail = act && (i == size - 1 || next != c[q]);
        const uint64_t mask = __ballot(head);
        const int r = runs + __popcll(mask & below) - 1;
        if (head) {
            bsr_columns[out0 + r] = c[q];
            beg_of[r] = i;
        }
        if (summed_offsets && tail)
            summed_offsets[out0 + r] = base + i + 1;
        if (act)
            idx_of[i] = (int)(uint32_t)k[q];
        runs += __popcll(mask);
    }
    return runs;
}

// Loads of one run sum issued ahead of its adds.
constexpr int kRunBatch = 8;

// Medium rows (sorted in memory by sort_medium_rows_kernel), one wave per
// listed row: columns, optionally inclusive run ends, and optionally the
// block sums.  The run starts and sorted triplet indices go through the
// wave's LDS.  A sum adds its run's blocks in ascending triplet order, the
// first as the initial value; a lane works on two (block, scalar) sums at
// once and issues kRunBatch loads of each before adding.
template <typename S, int BS>
__global__ __launch_bounds__(kWave * kRowsPerBlock) void finalize_medium_rows_kernel(
    RowRanges rr, const uint64_t* __restrict__ keys, const int* __restrict__ bsr_offsets,
    int* __restrict__ bsr_columns, int* __restrict__ summed_offsets, const S* __restrict__ tpl_values,
    int block_size, S* __restrict__ bsr_values, const int* __restrict__ medium_rows, const int* __restrict__ counters
)
{
    __shared__ int s_idx[kRowsPerBlock][kMediumCap];
    __shared__ int s_beg[kRowsPerBlock][kMediumCap + 1];
    const int bs = BS > 0 ? BS : block_size;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = (int)(threadIdx.x / kWave);
    int* idx_of = s_idx[wv];
    int* beg_of = s_beg[wv];
    const int nmed = counters[1];
    const int nwaves = gridDim.x * kRowsPerBlock;
    for (int li = blockIdx.x * kRowsPerBlock + wv; li < nmed; li += nwaves) {
        const int row = medium_rows[li];
        const int base = rr.base(row);
        const int size = rr.size(row);
        const int out0 = bsr_offsets[row];
        const int runs =
            size <= 2 * kWave
                ? medium_row_runs<2>(keys, base, size, out0, lane, bsr_columns, summed_offsets, idx_of, beg_of)
            : size <= 4 * kWave
                ? medium_row_runs<4>(keys, base, size, out0, lane, bsr_columns, summed_offsets, idx_of, beg_of)
                : medium_row_runs<kMediumRegs>(keys, base, size, out0, lane, bsr_columns, summed_offsets, idx_of,
                                               beg_of);
        if (bsr_values) {
            if (lane == 0)
                beg_of[runs] = size;
            wave_lds_sync();
            const int ntask = runs * bs;
            for (int t0 = lane; t0 < ntask; t0 += 2 * kWave) {
                int beg[2], end[2], sc[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int t = t0 + h * kWave;
                    const int u = t < ntask ? t / bs : 0;
                    sc[h] = t - u * bs;
                    beg[h] = t < ntask ? beg_of[u] : 0;
                    end[h] = t < ntask ? beg_of[u + 1] : 0;
                }
                S acc[2] = {S(0), S(0)};
                for (int o = 0; beg[0] + o < end[0] || beg[1] + o < end[1]; o += kRunBatch) {
                    S v[2][kRunBatch];
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int q = 0; q < kRunBatch; ++q)
                            if (beg[h] + o + q < end[h])
                                v[h][q] = tpl_values[(size_t)idx_of[beg[h] + o + q] * bs + sc[h]];
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int q = 0; q < kRunBatch; ++q)
                            if (beg[h] + o + q < end[h])
                                acc[h] = (o + q == 0) ? v[h][q] : acc[h] + v[h][q];
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    if (t0 + h * kWave < ntask)
                        bsr_values[(size_t)(out0 + (t0 + h * kWave) / bs) * bs + sc[h]] = acc[h];
            }
            wave_lds_sync();
        }
    }
}

// Rows above kMediumCap: one workgroup each (grid-stride over the row list).
__global__ __launch_bounds__(kBlock) void sort_large_rows_kernel(
    RowRanges rr, uint64_t* keys, uint64_t* alt, int* summed_indices, int* urow, const int* large_rows,
    const int* counters
)
{
    __shared__ uint64_t s[kLargeCap];
    __shared__ int count;
    const int nlarge = counters[0];
    const int tid = threadIdx.x;
    for (int li = blockIdx.x; li < nlarge; li += gridDim.x) {
        const int row = large_rows[li];
        const int base = rr.base(row);
        const int size = rr.size(row);
        if (tid == 0)
            count = 0;
        if (size <= kLargeCap) {
            const int n = pow2_at_least(size, 2);
            for (int k = tid; k < n; k += kBlock)
                s[k] = k < size ? keys[base + k] : kPadKey;
            __syncthreads();
            lds_bitonic(s, n, tid, kBlock);
            emit_sorted_row(s, size, base, row, tid, kBlock, keys, summed_indices, urow, &count);
            __syncthreads();
            continue;
        }
        // sort chunks of kLargeCap in LDS
        for (int c0 = 0; c0 < size; c0 += kLargeCap) {
            const int csize = size - c0 < kLargeCap ? size - c0 : kLargeCap;
            const int n = pow2_at_least(csize, 2);
            for (int k = tid; k < n; k += kBlock)
                s[k] = k < csize ? keys[base + c0 + k] : kPadKey;
            __syncthreads();
            lds_bitonic(s, n, tid, kBlock);
            for (int k = tid; k < csize; k += kBlock)
                store_agent(&keys[base + c0 + k], s[k]);
            __threadfence();
            __syncthreads();
        }
        // merge sorted runs by rank (keys are unique: the triplet index is in the low bits)
        uint64_t* src = keys + base;
        uint64_t* dst = alt + base;
        for (int w = kLargeCap; w < size; w <<= 1) {
            for (int p = tid; p < size; p += kBlock) {
                const int run0 = (p / (2 * w)) * (2 * w);
                const int a0 = run0, a1 = min(run0 + w, size);
                const int b0 = a1, b1 = min(run0 + 2 * w, size);
                const uint64_t key = load_agent(&src[p]);
                const bool in_a = p < a1;
                int lo = in_a ? b0 : a0;
                int hi = in_a ? b1 : a1;
                const int other0 = lo;
                while (lo < hi) {
                    const int mid = lo + (hi - lo) / 2;
                    if (load_agent(&src[mid]) < key)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                const int rank_other = lo - other0;
                const int own = in_a ? p - a0 : p - b0;
                store_agent(&dst[run0 + own + rank_other], key);
            }
            __threadfence();
            __syncthreads();
            uint64_t* t = src;
            src = dst;
            dst = t;
        }
        int heads = 0;
        for (int k = tid; k < size; k += kBlock) {
            const uint64_t key = load_agent(&src[k]);
            if (src != keys + base)
                keys[base + k] = key;
            summed_indices[base + k] = (int)(uint32_t)key;
            heads += (k == 0 || (load_agent(&src[k - 1]) >> 32) != (key >> 32)) ? 1 : 0;
        }
        atomicAdd(&count, heads);
        __syncthreads();
        if (tid == 0)
            urow[row] = count;
        __threadfence();
        __syncthreads();
    }
}

// Rows above kMediumCap keys (sorted in memory by the LDS / merge sort), one
// wave per listed row (kRowsPerBlock per workgroup): columns and inclusive run
// ends of the row's unique blocks.
__global__ __launch_bounds__(kWave * kRowsPerBlock) void finalize_large_rows_kernel(
    RowRanges rr, const uint64_t* __restrict__ keys, const int* __restrict__ bsr_offsets,
    int* __restrict__ bsr_columns, int* __restrict__ summed_offsets, const int* __restrict__ large_rows,
    const int* __restrict__ counters
)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int nlarge = counters[0];
    const int nwaves = gridDim.x * kRowsPerBlock;
    for (int li = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave); li < nlarge; li += nwaves) {
        const int row = large_rows[li];
        const int rs = rr.base(row);
        const int re = rs + rr.size(row);
        const int out0 = bsr_offsets[row];
        int runs = 0;
        for (int p0 = rs; p0 < re; p0 += kWave) {
            const int p = p0 + lane;
            const bool active = p < re;
            int col = -1, prev = -1, next = -1;
            if (active) {
                col = (int)(keys[p] >> 32);
                prev = p > rs ? (int)(keys[p - 1] >> 32) : -1;
                next = p + 1 < re ? (int)(keys[p + 1] >> 32) : -1;
            }
            const bool head = active && col != prev;
            const bool tail = active && col != next;
            const uint64_t mask = __ballot(head);
            const uint64_t below = lane == 63 ? ~uint64_t(0) : ((uint64_t(1) << (lane + 1)) - 1);
            const int run = runs + __popcll(mask & below) - 1;
            if (head)
                bsr_columns[out0 + run] = col;
            if (tail)
                summed_offsets[out0 + run] = p + 1;
            runs += __popcll(mask);
        }
    }
}

// Sequential run sums per (unique block, scalar) of the listed large rows,
// one wave per listed row (kRowsPerBlock per workgroup).  A sum adds its
// run's blocks in ascending triplet order, the first as the initial value;
// the sorted triplet indices and then the values are loaded kRunBatch ahead.
template <typename S>
__global__ __launch_bounds__(kWave * kRowsPerBlock) void accumulate_large_rows_kernel(
    RowRanges rr, const int* __restrict__ bsr_offsets, const int* __restrict__ summed_offsets,
    const int* __restrict__ summed_indices, const S* __restrict__ tpl_values, int block_size,
    S* __restrict__ bsr_values, const int* __restrict__ large_rows, const int* __restrict__ counters
)
{
    const int nlarge = counters[0];
    const int nwaves = gridDim.x * kRowsPerBlock;
    const int lane = threadIdx.x & (kWave - 1);
    for (int li = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave); li < nlarge; li += nwaves) {
        const int row = large_rows[li];
        const int u0 = bsr_offsets[row];
        const int u1 = bsr_offsets[row + 1];
        const int64_t tasks = (int64_t)(u1 - u0) * block_size;
        for (int64_t t = lane; t < tasks; t += kWave) {
            const int u = u0 + (int)(t / block_size);
            const int s = (int)(t % block_size);
            const int beg = u == u0 ? rr.base(row) : summed_offsets[u - 1];
            const int end = summed_offsets[u];
            S acc = S(0);
            for (int k = beg; k < end; k += kRunBatch) {
                int iq[kRunBatch];
#pragma unroll
                for (int q = 0; q < kRunBatch; ++q)
                    iq[q] = k + q < end ? summed_indices[k + q] : 0;
                S v[kRunBatch];
#pragma unroll
                for (int q = 0; q < kRunBatch; ++q)
                    if (k + q < end)
                        v[q] = tpl_values[(size_t)iq[q] * block_size + s];
#pragma unroll
                for (int q = 0; q < kRunBatch; ++q)
                    if (k + q < end)
                        acc = (k + q == beg) ? v[q] : acc + v[q];
            }
            bsr_values[(size_t)u * block_size + s] = acc;
        }
    }
}

// Rows of at most kWave keys, one wave per row: the keys are sorted again in
// registers, then the row's columns, (optionally) its inclusive run ends and
// sorted triplet indices, and (optionally) its block sums are written.  A sum
// adds the run's blocks in ascending triplet order, the first one as the
// initial value; the first four blocks of the runs of four (block, scalar)
// pairs per lane are loaded together ahead of the adds.
template <typename S, int BS>
__global__ __launch_bounds__(kWave * kRowsPerBlock) void finalize_small_rows_kernel(
    int nrow, RowRanges rr, const uint64_t* __restrict__ keys,
    const int* __restrict__ bsr_offsets, int* __restrict__ bsr_columns, int* __restrict__ summed_offsets,
    int* __restrict__ summed_indices, const S* __restrict__ tpl_values, int block_size, S* __restrict__ bsr_values
)
{
    const int bs = BS > 0 ? BS : block_size;
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave);
    if (row >= nrow)
        return;
    const int base = rr.base(row);
    const int size = rr.size(row);
    if (size > kWave)
        return;
    const uint64_t key = load_sorted_row(keys, base, size, lane);
    const int col = (int)(key >> 32);
    const int idx = (int)(uint32_t)key;
    const int prev = __shfl_up(col, 1);
    const int next = __shfl_down(col, 1);
    const bool head = lane < size && (lane == 0 || prev != col);
    const bool tail = lane < size && (lane == size - 1 || next != col);
    const uint64_t heads = __ballot(head);
    const uint64_t below = lane == 63 ? ~uint64_t(0) : ((uint64_t(1) << (lane + 1)) - 1);
    const int run = __popcll(heads & below) - 1;
    const int out0 = bsr_offsets[row];
    if (head)
        bsr_columns[out0 + run] = col;
    if (summed_offsets && tail)
        summed_offsets[out0 + run] = base + lane + 1;
    if (summed_indices && lane < size)
        summed_indices[base + lane] = idx;
    if (!bsr_values)
        return;
    // Run starts by run number (per-wave LDS; entry nu = size), so that any
    // lane finds the run of its (block, scalar) pair without a scan.
    __shared__ int run_start[kRowsPerBlock][kWave + 1];
    int* rs = run_start[threadIdx.x / kWave];
    const int nu = __popcll(heads);
    if (head)
        rs[run] = lane;
    if (lane == 0)
        rs[nu] = size;
    __builtin_amdgcn_wave_barrier();
    const uint64_t above = heads & ~below;
    int run_len = head ? (above ? __builtin_ctzll(above) : size) - lane : 0;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1)
        run_len = max(run_len, __shfl_xor(run_len, o));
    // kFinSlots (block, scalar) pairs per lane at a time, the first kFinDepth
    // blocks of each of their runs loaded together before any add (one
    // memory round trip for rows of up to 28 blocks); longer runs continue four deep.
    // Cross-lane reads (ds_bpermute) only in wave-uniform control flow.
    constexpr int kFinSlots = 4;
    constexpr int kFinDepth = 4;
    for (int t0 = 0; t0 < nu * bs; t0 += kFinSlots * kWave) {
        int beg[kFinSlots], end[kFinSlots], sc[kFinSlots], u[kFinSlots];
        bool act[kFinSlots];
#pragma unroll
        for (int h = 0; h < kFinSlots; ++h) {
            const int t = t0 + h * kWave + lane;
            act[h] = t < nu * bs;
            u[h] = act[h] ? t / bs : 0;
            sc[h] = t - u[h] * bs;
            beg[h] = rs[u[h]];
            end[h] = rs[u[h] + 1];
        }
        int iq[kFinSlots][kFinDepth];
#pragma unroll
        for (int h = 0; h < kFinSlots; ++h)
#pragma unroll
            for (int q = 0; q < kFinDepth; ++q)
                iq[h][q] = __shfl(idx, min(beg[h] + q, kWave - 1));
        S v[kFinSlots][kFinDepth];
#pragma unroll
        for (int h = 0; h < kFinSlots; ++h)
#pragma unroll
            for (int q = 0; q < kFinDepth; ++q)
                v[h][q] = (act[h] && beg[h] + q < end[h]) ? tpl_values[(size_t)iq[h][q] * bs + sc[h]] : S(0);
        S acc[kFinSlots];
#pragma unroll
        for (int h = 0; h < kFinSlots; ++h) {
            acc[h] = v[h][0];
#pragma unroll
            for (int q = 1; q < kFinDepth; ++q)
                if (beg[h] + q < end[h])
                    acc[h] += v[h][q];
        }
        for (int k = kFinDepth; k < run_len; k += 4) {
#pragma unroll
            for (int h = 0; h < kFinSlots; ++h) {
                int jq[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    jq[q] = __shfl(idx, min(beg[h] + k + q, kWave - 1));
                S w[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (act[h] && beg[h] + k + q < end[h])
                        w[q] = tpl_values[(size_t)jq[q] * bs + sc[h]];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (act[h] && beg[h] + k + q < end[h])
                        acc[h] += w[q];
            }
        }
#pragma unroll
        for (int h = 0; h < kFinSlots; ++h)
            if (act[h])
                bsr_values[(size_t)(out0 + u[h]) * bs + sc[h]] = acc[h];
    }
}

int launch_fill(int* p, int64_t n, int v, hipStream_t stream)
{
    if (n <= 0)
        return kOk;
    fill_i32_kernel<<<(unsigned)ceil_div(n, kBlock), kBlock, 0, stream>>>(p, n, v);
    WA_LAUNCH_CHECK();
    return kOk;
}

}  // namespace

size_t triplets_workspace_bytes(int nrow, int T)
{
    size_t needed = 0;
    carve_ws(nullptr, 0, nrow, T, true, &needed);
    return needed + 256;
}


// Grid of the grid-strided medium-row kernels (one wave per listed row): the
// list length is only known on the device, so the grid covers the worst case
// up to 16 workgroups per compute unit.
unsigned medium_grid(int nrow)
{
    const int64_t cap = (int64_t)device_cu_count() * 16;
    const int64_t need = ceil_div(nrow, kRowsPerBlock);
    return (unsigned)std::max<int64_t>(1, std::min(need, cap));
}

// Sort the keys of every row (rows given by rr) by (col, triplet index) and
// count its unique columns into urow[0..nrow).  Rows of more than kWave keys
// are listed in w.medium_rows (up to kMediumCap keys, count in
// w.counters[1]) or w.large_rows (count in w.counters[0]) and sorted in place
// in memory, with their sorted triplet indices in summed_indices.  Smaller
// rows are sorted in registers and, unless write_back, left unsorted in
// memory (the finalize pass sorts them again).
int sort_rows(int nrow, RowRanges rr, const TripletWs& w, int* summed_indices, int* urow, bool write_back,
              hipStream_t stream, bool counters_zeroed = false)
{
    if (!counters_zeroed)
        WA_TRY(launch_fill(w.counters, kListCounters, 0, stream));
    if (nrow > 0) {
        classify_rows_kernel<<<(unsigned)ceil_div(nrow, kBlock), kBlock, 0, stream>>>(
            nrow, rr, w.large_rows, w.medium_rows, w.counters
        );
        WA_LAUNCH_CHECK();
        count_small_rows_kernel<<<(unsigned)ceil_div(nrow, kRowsPerBlock), kWave * kRowsPerBlock, 0, stream>>>(
            nrow, rr, w.keys, summed_indices, write_back ? 1 : 0, urow
        );
        WA_LAUNCH_CHECK();
        sort_medium_rows_kernel<<<medium_grid(nrow), kWave * kRowsPerBlock, 0, stream>>>(
            rr, w.keys, summed_indices, urow, w.medium_rows, w.counters
        );
        WA_LAUNCH_CHECK();
        const int large_grid = device_cu_count() * 4;  // 32 KB of LDS each: up to 5 per compute unit
        sort_large_rows_kernel<<<(unsigned)large_grid, kBlock, 0, stream>>>(
            rr, w.keys, w.alt, summed_indices, urow, w.large_rows, w.counters
        );
        WA_LAUNCH_CHECK();
    }
    return kOk;
}

// Steps 5-7 for rows sorted by sort_rows: exclusive scan of the unique counts
// (urow, in bsr_offsets) into bsr_offsets, then columns, run ends / sorted
// indices (when the caller asked for them: own_summed false) and value sums.
int finalize_rows(int nrow, RowRanges rr, const TripletWs& w, bool own_summed, int* summed_offsets,
                  int* summed_indices, int* bsr_offsets, int* bsr_columns, const void* tpl_values, int scalar_size,
                  int block_size, void* bsr_values, int* bsr_nnz_host, hipStream_t stream)
{
    const bool sum_values = tpl_values && bsr_values;
    WA_TRY(exclusive_scan_i32(bsr_offsets, bsr_offsets, nrow, w.scan_ws, stream));
    if (bsr_nnz_host)
        WA_HIP(hipMemcpyAsync(bsr_nnz_host, bsr_offsets + nrow, sizeof(int), hipMemcpyDeviceToHost, stream));
    if (nrow == 0)
        return kOk;
    // rows of <= kWave keys: sorted again in registers, finalized and summed in one pass
    const unsigned grid = (unsigned)ceil_div(nrow, kRowsPerBlock);
    const int block = kWave * kRowsPerBlock;
    int* so = own_summed ? nullptr : summed_offsets;
    int* si = own_summed ? nullptr : summed_indices;
    if (sum_values && scalar_size == 4 && block_size == 9)
        finalize_small_rows_kernel<float, 9><<<grid, block, 0, stream>>>(
            nrow, rr, w.keys, bsr_offsets, bsr_columns, so, si, (const float*)tpl_values, block_size,
            (float*)bsr_values);
    else if (sum_values && scalar_size == 4 && block_size == 1)
        finalize_small_rows_kernel<float, 1><<<grid, block, 0, stream>>>(
            nrow, rr, w.keys, bsr_offsets, bsr_columns, so, si, (const float*)tpl_values, block_size,
            (float*)bsr_values);
    else if (sum_values && scalar_size == 4)
        finalize_small_rows_kernel<float, 0><<<grid, block, 0, stream>>>(
            nrow, rr, w.keys, bsr_offsets, bsr_columns, so, si, (const float*)tpl_values, block_size,
            (float*)bsr_values);
    else if (sum_values)
        finalize_small_rows_kernel<double, 0><<<grid, block, 0, stream>>>(
            nrow, rr, w.keys, bsr_offsets, bsr_columns, so, si, (const double*)tpl_values, block_size,
            (double*)bsr_values);
    else
        finalize_small_rows_kernel<float, 1><<<grid, block, 0, stream>>>(
            nrow, rr, w.keys, bsr_offsets, bsr_columns, so, si, nullptr, block_size, nullptr);
    WA_LAUNCH_CHECK();
    // rows of kWave + 1 .. kMediumCap keys (sorted in memory), one wave per listed row
    const unsigned mgrid = medium_grid(nrow);
    if (sum_values && scalar_size == 4 && block_size == 9)
        finalize_medium_rows_kernel<float, 9><<<mgrid, block, 0, stream>>>(
            rr, w.keys, bsr_offsets, bsr_columns, so, (const float*)tpl_values, block_size, (float*)bsr_values,
            w.medium_rows, w.counters);
    else if (sum_values && scalar_size == 4 && block_size == 1)
        finalize_medium_rows_kernel<float, 1><<<mgrid, block, 0, stream>>>(
            rr, w.keys, bsr_offsets, bsr_columns, so, (const float*)tpl_values, block_size, (float*)bsr_values,
            w.medium_rows, w.counters);
    else if (sum_values && scalar_size == 4)
        finalize_medium_rows_kernel<float, 0><<<mgrid, block, 0, stream>>>(
            rr, w.keys, bsr_offsets, bsr_columns, so, (const float*)tpl_values, block_size, (float*)bsr_values,
            w.medium_rows, w.counters);
    else if (sum_values)
        finalize_medium_rows_kernel<double, 0><<<mgrid, block, 0, stream>>>(
            rr, w.keys, bsr_offsets, bsr_columns, so, (const double*)tpl_values, block_size, (double*)bsr_values,
            w.medium_rows, w.counters);
    else
        finalize_medium_rows_kernel<float, 1><<<mgrid, block, 0, stream>>>(
            rr, w.keys, bsr_offsets, bsr_columns, so, nullptr, block_size, nullptr, w.medium_rows, w.counters);
    WA_LAUNCH_CHECK();
    // rows of more than kMediumCap keys (sorted in memory), one wave per listed row
    finalize_large_rows_kernel<<<mgrid, block, 0, stream>>>(
        rr, w.keys, bsr_offsets, bsr_columns, summed_offsets, w.large_rows, w.counters);
    WA_LAUNCH_CHECK();
    if (sum_values) {
        if (scalar_size == 4)
            accumulate_large_rows_kernel<float><<<mgrid, block, 0, stream>>>(
                rr, bsr_offsets, summed_offsets, summed_indices, (const float*)tpl_values, block_size,
                (float*)bsr_values, w.large_rows, w.counters);
        else
            accumulate_large_rows_kernel<double><<<mgrid, block, 0, stream>>>(
                rr, bsr_offsets, summed_offsets, summed_indices, (const double*)tpl_values, block_size,
                (double*)bsr_values, w.large_rows, w.counters);
        WA_LAUNCH_CHECK();
    }
    return kOk;
}

// Steps 1-4: bucket the valid triplets by row and sort every bucket by
// (col, triplet index); the unique column count of every row goes to
// urow[0..nrow).  Rows of more than kWave keys are listed in w.medium_rows /
// w.large_rows (counts in w.counters[1] / [0]) and left sorted in place at
// w.keys[w.row_start[r] ...] with their sorted triplet indices in
// summed_indices.  Smaller rows are sorted in registers and, unless
// write_back, left unsorted in memory (the finalize pass sorts them again).
int bucket_sort_rows(
    int block_size, int scalar_size, int nrow, int ncol, int T, const int* tpl_count, const int* rows, const int* cols,
    const void* tpl_values, uint64_t zero_mask, const int* mo, const int* mrc, const int* mc, const TripletWs& w,
    int* summed_indices, int* urow, bool write_back, hipStream_t stream
)
{
    const bool prune = tpl_values && zero_mask != 0;
    if (prune)
        WA_REQUIRE(scalar_size == 1 || scalar_size == 2 || scalar_size == 4 || scalar_size == 8,
                   "bsr_from_triplets: unsupported scalar size %d", scalar_size);
    if (w.part.tiled) {
        // coarse buckets through per-tile LDS histograms (no global atomics),
        // then each bucket's rows in LDS
        const Partition& pt = w.part;
        const unsigned tiles = (unsigned)pt.ntiles;
        switch (prune ? scalar_size : 0) {
        case 1:
            tile_classify_kernel<uint8_t><<<tiles, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint8_t*)tpl_values, block_size, (uint8_t)zero_mask, mo,
                mrc, mc, pt, w.trow, w.tile_offsets);
            break;
        case 2:
            tile_classify_kernel<uint16_t><<<tiles, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint16_t*)tpl_values, block_size, (uint16_t)zero_mask,
                mo, mrc, mc, pt, w.trow, w.tile_offsets);
            break;
        case 4:
            tile_classify_kernel<uint32_t><<<tiles, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint32_t*)tpl_values, block_size, (uint32_t)zero_mask,
                mo, mrc, mc, pt, w.trow, w.tile_offsets);
            break;
        case 8:
            tile_classify_kernel<uint64_t><<<tiles, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint64_t*)tpl_values, block_size, (uint64_t)zero_mask,
                mo, mrc, mc, pt, w.trow, w.tile_offsets);
            break;
        default:
            tile_classify_kernel<uint32_t><<<tiles, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, nullptr, block_size, 0u, mo, mrc, mc, pt, w.trow, w.tile_offsets);
            break;
        }
        WA_LAUNCH_CHECK();
        WA_TRY(exclusive_scan_i32(w.tile_offsets, w.tile_offsets, (int64_t)pt.nb * pt.ntiles, w.scan_ws, stream));
        // both scatters also zero the row histogram (w.hist, nrow + 1 entries)
        if (pt.staged)
            tile_scatter_staged_kernel<<<tiles, kStagedThreads, 0, stream>>>(T, w.trow, cols, pt, w.tile_offsets, w.alt,
                                                                             w.brow, w.hist, (int64_t)nrow + 1);
        else
            tile_scatter_kernel<<<tiles, kBlock, 0, stream>>>(T, w.trow, cols, pt, w.tile_offsets, w.alt, w.brow, w.hist,
                                                              (int64_t)nrow + 1);
        WA_LAUNCH_CHECK();
        // rows within the buckets: chunked LDS histograms -> scan -> chunked scatter
        const int* total = w.tile_offsets + (size_t)pt.nb * pt.ntiles;
        const unsigned chunks = (unsigned)(ceil_div(ceil_div(T, kChunk), kXcds) * kXcds);
        chunk_row_hist_kernel<<<chunks, kChunkThreads, 0, stream>>>(nrow, pt, total, w.brow, w.hist);
        WA_LAUNCH_CHECK();
        // row starts; the histogram is left zeroed for the scatter's run reservations
        WA_TRY(exclusive_scan_i32_clear(w.hist, w.row_start, nrow, w.scan_ws, stream));
        const unsigned schunks = (unsigned)(ceil_div(ceil_div(T, kSChunk), kXcds) * kXcds);
        chunk_row_scatter_kernel<<<schunks, kChunkThreads, 0, stream>>>(nrow, pt, total, w.alt, w.brow, w.row_start,
                                                                        w.hist, w.keys, w.counters);
        WA_LAUNCH_CHECK();
        return sort_rows(nrow, RowRanges{w.row_start, nullptr}, w, summed_indices, urow, write_back, stream,
                         /*counters_zeroed=*/true);
    }
    WA_TRY(launch_fill(w.hist, (int64_t)nrow + 1, 0, stream));
    if (T > 0) {
        const unsigned grid = (unsigned)ceil_div(T, kBlock);
        switch (prune ? scalar_size : 0) {
        case 1:
            classify_kernel<uint8_t><<<grid, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint8_t*)tpl_values, block_size, (uint8_t)zero_mask, mo,
                mrc, mc, w.trow, w.hist
            );
            break;
        case 2:
            classify_kernel<uint16_t><<<grid, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint16_t*)tpl_values, block_size, (uint16_t)zero_mask, mo,
                mrc, mc, w.trow, w.hist
            );
            break;
        case 4:
            classify_kernel<uint32_t><<<grid, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint32_t*)tpl_values, block_size, (uint32_t)zero_mask, mo,
                mrc, mc, w.trow, w.hist
            );
            break;
        case 8:
            classify_kernel<uint64_t><<<grid, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, (const uint64_t*)tpl_values, block_size, (uint64_t)zero_mask, mo,
                mrc, mc, w.trow, w.hist
            );
            break;
        default:
            classify_kernel<uint32_t><<<grid, kBlock, 0, stream>>>(
                T, tpl_count, nrow, ncol, rows, cols, nullptr, block_size, 0u, mo, mrc, mc, w.trow, w.hist
            );
            break;
        }
        WA_LAUNCH_CHECK();
    }
    WA_TRY(exclusive_scan_i32(w.hist, w.row_start, nrow, w.scan_ws, stream));
    if (T > 0) {
        scatter_kernel<<<(unsigned)ceil_div(T, kBlock), kBlock, 0, stream>>>(
            T, w.trow, cols, w.row_start, w.hist, w.keys
        );
        WA_LAUNCH_CHECK();
    }
    return sort_rows(nrow, RowRanges{w.row_start, nullptr}, w, summed_indices, urow, write_back, stream);
}

// Triplets -> compact BSR.  `mo`, `mrc`, `mc` (offsets, row counts, columns of
// an existing topology, or null) restrict the result to blocks of that
// topology; they may alias the outputs (the reference's masked mode), as they
// are read before the outputs are written.
int bsr_from_triplets(
    int block_size, int scalar_size, int nrow, int ncol, int T, const int* tpl_count, const int* rows, const int* cols,
    const void* tpl_values, uint64_t zero_mask, const int* mo, const int* mrc, const int* mc, int* summed_offsets,
    int* summed_indices, int* bsr_offsets, int* bsr_columns, void* bsr_values, int* bsr_nnz_host, void* workspace,
    size_t workspace_bytes, hipStream_t stream
)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && T >= 0 && block_size > 0, "bsr_from_triplets: invalid sizes");
    WA_REQUIRE(bsr_offsets != nullptr, "bsr_from_triplets: null offsets");
    WA_REQUIRE(T == 0 || (rows && cols && bsr_columns), "bsr_from_triplets: null triplet or column arrays");
    const bool own_summed = !(summed_offsets && summed_indices);
    size_t needed = 0;
    TripletWs w = carve_ws(workspace, workspace_bytes, nrow, T, own_summed, &needed);
    WA_REQUIRE(workspace && needed <= workspace_bytes,
               "bsr_from_triplets: workspace too small (%zu < %zu bytes)", workspace_bytes, needed);
    if (own_summed) {
        summed_offsets = w.summed_offsets;
        summed_indices = w.summed_indices;
    }
    const bool sum_values = tpl_values && bsr_values;
    if (sum_values)
        WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_from_triplets: value accumulation needs fp32 or fp64");

    // unique counts land in bsr_offsets[0..nrow), scanned in place by finalize_rows
    WA_TRY(bucket_sort_rows(block_size, scalar_size, nrow, ncol, T, tpl_count, rows, cols, tpl_values, zero_mask, mo,
                            mrc, mc, w, summed_indices, bsr_offsets, false, stream));
    return finalize_rows(nrow, RowRanges{w.row_start, nullptr}, w, own_summed, summed_offsets, summed_indices,
                         bsr_offsets, bsr_columns, sum_values ? tpl_values : nullptr, scalar_size, block_size,
                         sum_values ? bsr_values : nullptr, bsr_nnz_host, stream);
}


// ---------------------------------------------------------------------------
// bsr_compress: sort and coalesce the active blocks of each row (reference
// bsr_compress, warp/_src/sparse.py:1571-1672; per-row semantics of
// bsr_compress_inplace_row, warp/native/sparse_util.h:82-146).  The active
// blocks become triplets (row, column, storage position) for the assembly
// above: a stable sort by (row, column, position) and sequential sums in
// storage order are exactly the reference's in-row insertion sort and
// left-to-right accumulation.
// ---------------------------------------------------------------------------

// rows[k] = r for every k in row r's active range [offsets[r], end), k < T;
// the range never passes the row's capacity end offsets[r+1] (as the
// reference's bsr_active_row_end clipping, warp/native/sparse.cpp:445-447)
__global__ __launch_bounds__(kWave) void active_rows_kernel(int nrow, const int* __restrict__ offsets,
                                                           const int* __restrict__ row_counts, int T,
                                                           int* __restrict__ rows)
{
    const int r = blockIdx.x;
    if (r >= nrow)
        return;
    const int beg = offsets[r];
    const int cap_end = offsets[r + 1];
    int end = row_counts ? beg + row_counts[r] : cap_end;
    end = end < cap_end ? end : cap_end;
    end = end < T ? end : T;
    for (int k = beg + (int)threadIdx.x; k < end; k += kWave)
        rows[k] = r;
}

// The active blocks of row r become keys (column << 32 | position) at
// keys[offsets[r] ...] in storage order (ballot compaction, a wave per row),
// rcount[r] their number.  Dropped, as by the triplet classification:
// columns outside [0, ncol) and, with values, all-zero blocks.  Positions
// beyond T or the row's capacity are never active.
template <typename U>
__global__ __launch_bounds__(kWave * kRowsPerBlock) void compress_keys_kernel(
    int nrow, int ncol, int T, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, const U* __restrict__ values, int block_size, U zero_mask,
    uint64_t* __restrict__ keys, int* __restrict__ rcount)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave);
    if (row >= nrow)
        return;
    const int beg = offsets[row];
    const int cap_end = offsets[row + 1];
    int end = row_counts ? beg + row_counts[row] : cap_end;
    end = end < cap_end ? end : cap_end;
    end = end < T ? end : T;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    int count = 0;
    for (int p0 = beg; p0 < end; p0 += kWave) {
        const int p = p0 + lane;
        bool valid = p < end;
        const int c = valid ? columns[p] : -1;
        valid = valid && c >= 0 && c < ncol;
        if (valid && values)
            valid = !block_is_zero(values + (size_t)p * block_size, block_size, zero_mask);
        const uint64_t mask = __ballot(valid);
        if (valid)
            keys[beg + count + __popcll(mask & below)] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)p;
        count += __popcll(mask);
    }
    if (lane == 0)
        rcount[row] = count;
}

// Compact result written back over the matrix: columns / values of the
// nnz = c_offsets[nrow] coalesced blocks, then column -1 and zero values up to
// position T (values moved as 32-bit words).
__global__ __launch_bounds__(kBlock) void compact_copy_kernel(int nrow, const int* __restrict__ c_offsets, int T,
                                                             int block_words, const int* __restrict__ c_columns,
                                                             const uint32_t* __restrict__ c_values,
                                                             int* __restrict__ columns, uint32_t* __restrict__ values)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int nnz = c_offsets[nrow];
    if (t < T)
        columns[t] = t < nnz ? c_columns[t] : -1;
    if (values && t < (int64_t)T * block_words)
        values[t] = t < (int64_t)nnz * block_words ? c_values[t] : 0u;
}

// Padded result: row r's compact blocks go to the start of its capacity
// range [offsets[r], offsets[r+1]) and row_counts[r] is set.  A row whose
// blocks do not fit gets row_counts[r] = 0 and raises *status to
// ROW_CAPACITY_EXCEEDED (reference _bsr_set_from_triplets_padded_scatter_
// compact_topology, warp/_src/sparse.py:1173-1208).  Values move as 32-bit words.
__global__ __launch_bounds__(kWave) void padded_scatter_kernel(int nrow, const int* __restrict__ offsets,
                                                              int* __restrict__ row_counts,
                                                              const int* __restrict__ c_offsets,
                                                              const int* __restrict__ c_columns,
                                                              const uint32_t* __restrict__ c_values, int block_words,
                                                              int* __restrict__ columns, uint32_t* __restrict__ values,
                                                              int* __restrict__ status)
{
    const int r = blockIdx.x;
    if (r >= nrow)
        return;
    const int src = c_offsets[r];
    const int cnt = c_offsets[r + 1] - src;
    const int dst = offsets[r];
    if (cnt > offsets[r + 1] - dst) {
        if (threadIdx.x == 0) {
            row_counts[r] = 0;
            if (status)
                atomicMax(status, kStatusRowCapacityExceeded);
        }
        return;
    }
    if (threadIdx.x == 0)
        row_counts[r] = cnt;
    for (int j = threadIdx.x; j < cnt; j += kWave)
        columns[dst + j] = c_columns[src + j];
    if (values)
        for (int b = threadIdx.x; b < cnt * block_words; b += kWave)
            values[(size_t)dst * block_words + b] = c_values[(size_t)src * block_words + b];
}

// Masked result: every active block (r, c) of the existing topology takes the
// compact block (r, c) when there is one; a repeated column keeps only its
// first slot, like the reference's bsr_block_index lookup
// (_bsr_set_from_triplets_masked_values, warp/_src/sparse.py:1145-1170).
// Values were cleared beforehand.
__global__ __launch_bounds__(kWave) void masked_scatter_kernel(int nrow, const int* __restrict__ offsets,
                                                              const int* __restrict__ row_counts,
                                                              const int* __restrict__ columns,
                                                              const int* __restrict__ c_offsets,
                                                              const int* __restrict__ c_columns,
                                                              const uint32_t* __restrict__ c_values, int block_words,
                                                              uint32_t* __restrict__ values)
{
    const int r = blockIdx.x;
    if (r >= nrow)
        return;
    const int beg = offsets[r];
    const int end = row_counts ? beg + row_counts[r] : offsets[r + 1];
    const int c0 = c_offsets[r];
    const int c1 = c_offsets[r + 1];
    for (int k = beg + (int)threadIdx.x; k < end; k += kWave) {
        const int col = columns[k];
        if (k > beg && columns[k - 1] == col)
            continue;
        int lo = c0, hi = c1;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (c_columns[mid] < col)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < c1 && c_columns[lo] == col)
            for (int b = 0; b < block_words; ++b)
                values[(size_t)k * block_words + b] = c_values[(size_t)lo * block_words + b];
    }
}


// ---------------------------------------------------------------------------
// Transpose (reference wp_bsr_transpose_device, warp/native/sparse.cu:1400-1492,
// and bsr_set_transpose, warp/_src/sparse.py:2636-2833).  Every active block
// (r, c) at storage position p becomes the triplet (row c, column r, index p);
// the row bucketing and per-row sort of the assembly order each transposed row
// by (r, p), the same order as the reference's sort of (c, r) keys with the
// source block index as tie-break.  Nothing is merged: a repeated block stays
// repeated.
// ---------------------------------------------------------------------------

// Compact transposed topology: columns and source block of each sorted key.
__global__ __launch_bounds__(kBlock) void transpose_emit_kernel(int T, const int* __restrict__ total,
                                                               const uint64_t* __restrict__ keys,
                                                               int* __restrict__ t_columns, int* __restrict__ block_map)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= T || i >= *total)
        return;
    const uint64_t key = keys[i];
    t_columns[i] = (int)(key >> 32);
    block_map[i] = (int)(uint32_t)key;
}

// Padded transposed topology: row r's sorted keys go to the start of its
// capacity range, block_map is -1 on the rest of the range; an overflowing row
// is left empty and raises *status (reference bsr_transpose_scatter_padded,
// warp/native/sparse.cu:383-428).
__global__ __launch_bounds__(kWave) void transpose_padded_kernel(int nrow, const int* __restrict__ row_start,
                                                                const uint64_t* __restrict__ keys,
                                                                const int* __restrict__ t_offsets,
                                                                int* __restrict__ t_row_counts,
                                                                int* __restrict__ t_columns,
                                                                int* __restrict__ block_map, int* __restrict__ status)
{
    const int r = blockIdx.x;
    if (r >= nrow)
        return;
    const int src = row_start[r];
    const int cnt = row_start[r + 1] - src;
    const int beg = t_offsets[r];
    const int cap = t_offsets[r + 1] - beg;
    const bool fits = cnt <= cap;
    if (threadIdx.x == 0) {
        t_row_counts[r] = fits ? cnt : 0;
        if (!fits && status)
            atomicMax(status, kStatusRowCapacityExceeded);
    }
    const int used = fits ? cnt : 0;
    for (int j = threadIdx.x; j < cap; j += kWave) {
        if (j < used) {
            const uint64_t key = keys[src + j];
            t_columns[beg + j] = (int)(key >> 32);
            block_map[beg + j] = (int)(uint32_t)key;
        } else {
            block_map[beg + j] = -1;
        }
    }
}

// dst[b] = scale * transpose(src[block_map[b]]) for b < *total with a source
// block (reference _bsr_transpose_values, warp/_src/sparse.py:2604-2633); one
// thread per destination scalar.  Source blocks are br x bc.
template <typename S>
__global__ __launch_bounds__(kBlock) void transpose_values_kernel(int64_t count, const int* __restrict__ total, int br,
                                                                 int bc, const int* __restrict__ block_map,
                                                                 const S* __restrict__ src, S scale,
                                                                 S* __restrict__ dst)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int bs = br * bc;
    if (t >= count * bs)
        return;
    const int64_t b = t / bs;
    if (b >= *total)
        return;
    const int sb = block_map[b];
    if (sb < 0)
        return;
    const int e = (int)(t - b * bs);
    const int i = e / br;   // destination block row = source column
    const int j = e - i * br;
    dst[t] = src[(size_t)sb * bs + (size_t)j * bc + i] * scale;
}

// Masked transpose: each active destination block (r, c) takes
// scale * transpose of source block (c, r) when the source has it (a binary
// search in source row c); values were cleared beforehand.  One wave per
// destination row.
template <typename S>
__global__ __launch_bounds__(kWave) void transpose_masked_values_kernel(
    int t_nrow, const int* __restrict__ t_offsets, const int* __restrict__ t_row_counts,
    const int* __restrict__ t_columns, int nrow, const int* __restrict__ offsets, const int* __restrict__ row_counts,
    const int* __restrict__ columns, int br, int bc, const S* __restrict__ src, S scale, S* __restrict__ dst)
{
    const int r = blockIdx.x;
    if (r >= t_nrow)
        return;
    const int beg = t_offsets[r];
    const int end = t_row_counts ? beg + t_row_counts[r] : t_offsets[r + 1];
    const int bs = br * bc;
    for (int k = beg + (int)threadIdx.x; k < end; k += kWave) {
        const int c = t_columns[k];
        if (c < 0 || c >= nrow)
            continue;
        const int s0 = offsets[c];
        const int s1 = row_counts ? s0 + row_counts[c] : offsets[c + 1];
        int lo = s0, hi = s1;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (columns[mid] < r)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo >= s1 || columns[lo] != r)
            continue;
        for (int e = 0; e < bs; ++e) {
            const int i = e / br;
            const int j = e - i * br;
            dst[(size_t)k * bs + e] = src[(size_t)lo * bs + (size_t)j * bc + i] * scale;
        }
    }
}

// Workspace of the compress / padded / masked / transpose drivers: source row
// of every storage position (compress, transpose), a compact staging result,
// and the assembly's own workspace.
struct StageWs {
    int* rows;
    int* c_offsets;
    int* c_columns;
    uint32_t* c_values;
    int* urow;
    void* tws;
    size_t tws_bytes;
};

StageWs carve_stage(void* base, size_t bytes, int nrow, int bucket_rows, int T, int block_bytes, bool need_rows,
                    size_t* needed)
{
    Carve c(base, bytes);
    StageWs w;
    w.rows = need_rows ? c.take<int>((size_t)T) : nullptr;
    w.c_offsets = c.take<int>((size_t)nrow + 1);
    w.c_columns = c.take<int>((size_t)T);
    w.c_values = block_bytes > 0 ? c.take<uint32_t>((size_t)T * (block_bytes / 4)) : nullptr;
    w.urow = c.take<int>((size_t)bucket_rows + 1);
    const size_t tneed = triplets_workspace_bytes(bucket_rows, T);
    w.tws = c.take<char>(tneed);
    w.tws_bytes = tneed;
    *needed = c.used;
    return w;
}

int launch_zero_words(void* p, int64_t words, hipStream_t stream)
{
    return launch_fill(static_cast<int*>(p), words, 0, stream);
}

// Source row of every active storage position, -1 elsewhere.
int mark_active_rows(int nrow, const int* offsets, const int* row_counts, int T, int* rows, hipStream_t s)
{
    WA_TRY(launch_fill(rows, T, -1, s));
    if (T > 0 && nrow > 0) {
        active_rows_kernel<<<(unsigned)nrow, kWave, 0, s>>>(nrow, offsets, row_counts, T, rows);
        WA_LAUNCH_CHECK();
    }
    return kOk;
}

int bsr_transpose(int nrow, int ncol, int T, const int* offsets, const int* row_counts, const int* columns,
                  int* t_offsets, int* t_row_counts, int* t_columns, int* block_map, int* status, void* ws,
                  size_t bytes, hipStream_t s)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && T >= 0, "bsr_transpose: invalid sizes");
    WA_REQUIRE(offsets && t_offsets && (T == 0 || (columns && t_columns && block_map)),
               "bsr_transpose: null topology arrays");
    size_t needed = 0;
    StageWs w = carve_stage(ws, bytes, 0, ncol, T, 0, true, &needed);
    WA_REQUIRE(ws && needed <= bytes, "bsr_transpose: workspace too small (%zu < %zu bytes)", bytes, needed);
    size_t tneed = 0;
    TripletWs tw = carve_ws(w.tws, w.tws_bytes, ncol, T, true, &tneed);
    WA_TRY(mark_active_rows(nrow, offsets, row_counts, T, w.rows, s));
    // triplet i = (row columns[i], column rows[i]); inactive positions have column -1 and are dropped
    WA_TRY(bucket_sort_rows(1, 0, ncol, nrow, T, nullptr, columns, w.rows, nullptr, 0, nullptr, nullptr, nullptr, tw,
                            tw.summed_indices, w.urow, true, s));
    if (t_row_counts == nullptr) {
        WA_HIP(hipMemcpyAsync(t_offsets, tw.row_start, sizeof(int) * ((size_t)ncol + 1), hipMemcpyDeviceToDevice, s));
        if (T > 0) {
            transpose_emit_kernel<<<(unsigned)ceil_div(T, kBlock), kBlock, 0, s>>>(T, tw.row_start + ncol, tw.keys,
                                                                                   t_columns, block_map);
            WA_LAUNCH_CHECK();
        }
    } else if (ncol > 0) {
        transpose_padded_kernel<<<(unsigned)ncol, kWave, 0, s>>>(ncol, tw.row_start, tw.keys, t_offsets, t_row_counts,
                                                                 t_columns, block_map, status);
        WA_LAUNCH_CHECK();
    }
    return kOk;
}

size_t transpose_workspace_bytes(int ncol, int T)
{
    size_t needed = 0;
    carve_stage(nullptr, 0, 0, ncol, T, 0, true, &needed);
    return needed + 256;
}

}  // namespace wa

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

WA_API size_t warp_amd_bsr_from_triplets_workspace_size(int row_count, int nnz)
{
    return wa::triplets_workspace_bytes(row_count, nnz);
}


WA_API int warp_amd_bsr_from_triplets(
    int block_size, int scalar_size, int row_count, int col_count, int nnz, const int* tpl_nnz, const int* tpl_rows,
    const int* tpl_columns, const void* tpl_values, uint64_t scalar_zero_mask, int masked_topology,
    int* summed_block_offsets, int* summed_block_indices, int* bsr_offsets, const int* bsr_row_counts,
    int* bsr_columns, void* bsr_values, int* bsr_nnz_host, void* workspace, size_t workspace_bytes, void* stream
)
{
    return wa::bsr_from_triplets(
        block_size, scalar_size, row_count, col_count, nnz, tpl_nnz, tpl_rows, tpl_columns, tpl_values,
        scalar_zero_mask, masked_topology ? bsr_offsets : nullptr, masked_topology ? bsr_row_counts : nullptr,
        masked_topology ? bsr_columns : nullptr, summed_block_offsets, summed_block_indices, bsr_offsets, bsr_columns,
        bsr_values, bsr_nnz_host, workspace, workspace_bytes, wa::as_stream(stream)
    );
}

namespace {
thread_local void* g_current_stream = nullptr;
}

WA_API void warp_amd_set_current_stream(void* stream) { g_current_stream = stream; }

WA_API void* warp_amd_get_current_stream(void) { return g_current_stream; }

// Same name and arguments as the reference native entry point, so a caller of
// the reference library can bind this one unchanged.  Uses the stream set by
// warp_amd_set_current_stream (the reference uses its current CUDA stream),
// allocates its scratch with the stream-ordered allocator, and records
// `bsr_nnz_event` (a hipEvent_t) after the nnz copy when given.  Errors are
// reported through warp_amd_last_error() (the reference function returns void).
WA_API void wp_bsr_matrix_from_triplets_device(
    int block_size, int scalar_size, int row_count, int col_count, int nnz, const int* tpl_nnz, const int* tpl_rows,
    const int* tpl_columns, const void* tpl_values, uint64_t scalar_zero_mask, bool masked_topology,
    int* tpl_block_offsets, int* tpl_block_indices, int* bsr_offsets, const int* bsr_row_counts, int* bsr_columns,
    int* bsr_nnz, void* bsr_nnz_event
)
{
    hipStream_t stream = wa::as_stream(g_current_stream);
    const size_t bytes = wa::triplets_workspace_bytes(row_count, nnz);
    void* ws = nullptr;
    hipError_t err = hipMallocAsync(&ws, bytes, stream);
    if (err != hipSuccess) {
        wa::set_error("wp_bsr_matrix_from_triplets_device: hipMallocAsync failed: %s", hipGetErrorString(err));
        return;
    }
    wa::bsr_from_triplets(
        block_size, scalar_size, row_count, col_count, nnz, tpl_nnz, tpl_rows, tpl_columns, tpl_values,
        scalar_zero_mask, masked_topology ? bsr_offsets : nullptr, masked_topology ? bsr_row_counts : nullptr,
        masked_topology ? bsr_columns : nullptr, tpl_block_offsets, tpl_block_indices, bsr_offsets, bsr_columns,
        nullptr, bsr_nnz, ws, bytes, stream
    );
    if (bsr_nnz && bsr_nnz_event)
        (void)hipEventRecord(static_cast<hipEvent_t>(bsr_nnz_event), stream);
    (void)hipFreeAsync(ws, stream);
}

// ---------------------------------------------------------------------------
// bsr_compress
// ---------------------------------------------------------------------------

// Workspace of warp_amd_bsr_compress for nrow block rows, T storage
// positions and blocks of block_bytes bytes (0 for a topology-only call).
WA_API size_t warp_amd_bsr_compress_workspace_size(int nrow, int T, int block_bytes)
{
    size_t needed = 0;
    wa::carve_stage(nullptr, 0, nrow, nrow, T, block_bytes, true, &needed);
    return needed + 256;
}

// Sort and coalesce the active blocks of each row in place (replaces
// wp_bsr_compress_inplace_device / _host, warp/native/sparse.cpp:521-607,
// and bsr_compress, warp/_src/sparse.py:1571-1672).  Active blocks are the
// positions [offsets[r], offsets[r] + row_counts[r]) (row_counts NULL:
// [offsets[r], offsets[r+1])) below T; negative columns are dropped, and with
// prune_numerical_zeros so are blocks whose scalars all satisfy
// (bits & zero_mask) == 0.  Duplicates are summed left to right in storage
// order (values NULL: topology only).  make_compact = 1 packs the rows
// (offsets rewritten, positions beyond the new nnz get column -1 and zero
// values, *nnz_host = nnz when given); 0 keeps offsets and writes
// row_counts.  Only fp32 / fp64 values can be summed.
WA_API int warp_amd_bsr_compress(int block_size, int scalar_size, int nrow, int ncol, int T, int prune_numerical_zeros,
                                 uint64_t zero_mask, int make_compact, int* offsets, int* row_counts, int* columns,
                                 void* values, int* nnz_host, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && T >= 0 && block_size > 0, "bsr_compress: invalid sizes");
    WA_REQUIRE(offsets != nullptr && (T == 0 || columns != nullptr), "bsr_compress: null topology arrays");
    WA_REQUIRE(make_compact || row_counts != nullptr, "bsr_compress: padded output needs row_counts");
    WA_REQUIRE(!values || scalar_size == 4 || scalar_size == 8, "bsr_compress: values must be fp32 or fp64");
    hipStream_t s = wa::as_stream(stream);
    const int block_bytes = values ? block_size * scalar_size : 0;
    const int block_words = block_bytes / 4;
    size_t needed = 0;
    wa::StageWs w = wa::carve_stage(workspace, workspace_bytes, nrow, nrow, T, block_bytes, true, &needed);
    WA_REQUIRE(workspace && needed <= workspace_bytes, "bsr_compress: workspace too small (%zu < %zu bytes)",
               workspace_bytes, needed);
    if (nrow == 0) {
        if (make_compact && nnz_host)
            *nnz_host = 0;
        return wa::kOk;
    }
    // the candidates are already grouped by row: keys are built in place per row,
    // with no row bucketing, then sorted, counted and summed by the assembly core
    size_t tneed = 0;
    wa::TripletWs tw = wa::carve_ws(w.tws, w.tws_bytes, nrow, T, true, &tneed);
    int* rcount = w.urow;
    const bool prune = prune_numerical_zeros && values && zero_mask != 0;
    const unsigned grid = (unsigned)wa::ceil_div(nrow, wa::kRowsPerBlock);
    const int block = wa::kWave * wa::kRowsPerBlock;
    if (prune && scalar_size == 8)
        wa::compress_keys_kernel<uint64_t><<<grid, block, 0, s>>>(nrow, ncol, T, offsets, row_counts, columns,
                                                                  static_cast<const uint64_t*>(values), block_size,
                                                                  zero_mask, tw.keys, rcount);
    else
        wa::compress_keys_kernel<uint32_t><<<grid, block, 0, s>>>(
            nrow, ncol, T, offsets, row_counts, columns, prune ? static_cast<const uint32_t*>(values) : nullptr,
            block_size, (uint32_t)zero_mask, tw.keys, rcount);
    WA_LAUNCH_CHECK();
    const wa::RowRanges rr{offsets, rcount};
    WA_TRY(wa::sort_rows(nrow, rr, tw, tw.summed_indices, w.c_offsets, false, s));
    WA_TRY(wa::finalize_rows(nrow, rr, tw, true, tw.summed_offsets, tw.summed_indices, w.c_offsets, w.c_columns,
                             values, scalar_size, block_size, values ? (void*)w.c_values : nullptr,
                             make_compact ? nnz_host : nullptr, s));
    if (make_compact) {
        WA_HIP(hipMemcpyAsync(offsets, w.c_offsets, sizeof(int) * ((size_t)nrow + 1), hipMemcpyDeviceToDevice, s));
        if (T > 0) {
            const int64_t n = (int64_t)T * (block_words > 1 ? block_words : 1);
            wa::compact_copy_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, s>>>(
                nrow, w.c_offsets, T, block_words, w.c_columns, w.c_values, columns, static_cast<uint32_t*>(values));
            WA_LAUNCH_CHECK();
        }
    } else {
        wa::padded_scatter_kernel<<<(unsigned)nrow, wa::kWave, 0, s>>>(
            nrow, offsets, row_counts, w.c_offsets, w.c_columns, w.c_values, block_words, columns,
            static_cast<uint32_t*>(values), nullptr
        );
        WA_LAUNCH_CHECK();
    }
    return wa::kOk;
}

// ---------------------------------------------------------------------------
// Padded and masked triplet assembly
// ---------------------------------------------------------------------------

// Workspace of warp_amd_bsr_set_from_triplets_padded / _masked for nrow block
// rows, T triplets and blocks of block_bytes bytes (0: topology only).
WA_API size_t warp_amd_bsr_topology_workspace_size(int nrow, int T, int block_bytes)
{
    size_t needed = 0;
    wa::carve_stage(nullptr, 0, nrow, nrow, T, block_bytes, false, &needed);
    return needed + 256;
}

// Triplets into the existing row capacity of a padded matrix (reference
// bsr_set_from_triplets with topology="padded", warp/_src/sparse.py:1335-1420).
// offsets[0..nrow] is read as capacity and left unchanged; row r receives its
// sorted unique columns at offsets[r].. and row_counts[r] their number.  A row
// whose blocks exceed its capacity is left empty and raises *status (sticky) to
// ROW_CAPACITY_EXCEEDED.  With values, the first `capacity` blocks of `values`
// are cleared and the summed blocks written (duplicates summed in ascending
// triplet order, as the compact path).  Invalid and (zero_mask != 0) all-zero
// triplets are dropped as in warp_amd_bsr_from_triplets.
WA_API int warp_amd_bsr_set_from_triplets_padded(int block_size, int scalar_size, int nrow, int ncol, int T,
                                                 const int* tpl_count, const int* rows, const int* cols,
                                                 const void* tpl_values, uint64_t zero_mask, const int* offsets,
                                                 int* row_counts, int* columns, void* values, int64_t capacity,
                                                 int* status, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && T >= 0 && block_size > 0 && capacity >= 0, "bsr_set_from_triplets: invalid sizes");
    WA_REQUIRE(offsets && row_counts && (T == 0 || (rows && cols && columns)),
               "bsr_set_from_triplets: null topology arrays");
    const bool with_values = tpl_values && values;
    WA_REQUIRE(!with_values || scalar_size == 4 || scalar_size == 8,
               "bsr_set_from_triplets: values must be fp32 or fp64");
    hipStream_t s = wa::as_stream(stream);
    const int block_bytes = with_values ? block_size * scalar_size : 0;
    const int block_words = block_bytes / 4;
    size_t needed = 0;
    wa::StageWs w = wa::carve_stage(workspace, workspace_bytes, nrow, nrow, T, block_bytes, false, &needed);
    WA_REQUIRE(workspace && needed <= workspace_bytes, "bsr_set_from_triplets: workspace too small (%zu < %zu bytes)",
               workspace_bytes, needed);
    if (nrow == 0)
        return wa::kOk;
    WA_TRY(wa::bsr_from_triplets(block_size, scalar_size, nrow, ncol, T, tpl_count, rows, cols, tpl_values, zero_mask,
                                 nullptr, nullptr, nullptr, nullptr, nullptr, w.c_offsets, w.c_columns,
                                 with_values ? (void*)w.c_values : nullptr, nullptr, w.tws, w.tws_bytes, s));
    if (with_values)
        WA_TRY(wa::launch_zero_words(values, capacity * block_words, s));
    wa::padded_scatter_kernel<<<(unsigned)nrow, wa::kWave, 0, s>>>(
        nrow, offsets, row_counts, w.c_offsets, w.c_columns, w.c_values, block_words, columns,
        with_values ? static_cast<uint32_t*>(values) : nullptr, status
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Triplets into the existing topology of `offsets` / `row_counts` (NULL:
// compact) / `columns`, which is left unchanged (reference
// bsr_set_from_triplets with topology="masked", warp/_src/sparse.py:1422-1441):
// the first `capacity` blocks of `values` are cleared, and every active block
// (r, c) receives the sum of the valid triplets at (r, c), in ascending
// triplet order (the reference adds them with atomics, in no fixed order).
// Triplets outside the topology are dropped; none is pruned as zero.
WA_API int warp_amd_bsr_set_from_triplets_masked(int block_size, int scalar_size, int nrow, int ncol, int T,
                                                 const int* tpl_count, const int* rows, const int* cols,
                                                 const void* tpl_values, const int* offsets, const int* row_counts,
                                                 const int* columns, void* values, int64_t capacity, void* workspace,
                                                 size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(nrow >= 0 && ncol >= 0 && T >= 0 && block_size > 0 && capacity >= 0, "bsr_set_from_triplets: invalid sizes");
    WA_REQUIRE(offsets && values && tpl_values && (T == 0 || (rows && cols && columns)),
               "bsr_set_from_triplets: null arrays");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_set_from_triplets: values must be fp32 or fp64");
    hipStream_t s = wa::as_stream(stream);
    const int block_bytes = block_size * scalar_size;
    const int block_words = block_bytes / 4;
    size_t needed = 0;
    wa::StageWs w = wa::carve_stage(workspace, workspace_bytes, nrow, nrow, T, block_bytes, false, &needed);
    WA_REQUIRE(workspace && needed <= workspace_bytes, "bsr_set_from_triplets: workspace too small (%zu < %zu bytes)",
               workspace_bytes, needed);
    WA_TRY(wa::launch_zero_words(values, capacity * block_words, s));
    if (nrow == 0)
        return wa::kOk;
    WA_TRY(wa::bsr_from_triplets(block_size, scalar_size, nrow, ncol, T, tpl_count, rows, cols, tpl_values, 0, offsets,
                                 row_counts, columns, nullptr, nullptr, w.c_offsets, w.c_columns, w.c_values, nullptr,
                                 w.tws, w.tws_bytes, s));
    wa::masked_scatter_kernel<<<(unsigned)nrow, wa::kWave, 0, s>>>(nrow, offsets, row_counts, columns, w.c_offsets,
                                                                   w.c_columns, w.c_values, block_words,
                                                                   static_cast<uint32_t*>(values));
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// ---------------------------------------------------------------------------
// Transpose
// ---------------------------------------------------------------------------

// Workspace of warp_amd_bsr_transpose for a source with ncol block columns and
// T storage positions.
WA_API size_t warp_amd_bsr_transpose_workspace_size(int ncol, int T) { return wa::transpose_workspace_bytes(ncol, T); }

// Topology of the transpose of an nrow x ncol block matrix (replaces
// wp_bsr_transpose_device / _host, warp/native/sparse.cu:1400-1492,
// sparse.cpp:415-519).  The active blocks of the source (positions below T)
// are reordered by (column, row, position).  t_row_counts NULL: compact
// output, t_offsets[0..ncol] and t_columns / block_map[0..nnz) written.
// Otherwise padded output: t_offsets is read as row capacity; t_row_counts,
// t_columns and block_map over the capacity are written (block_map -1 on
// unused slots); a row that does not fit is left empty and raises *status.
// block_map[b] is the source position of transposed block b.
WA_API int warp_amd_bsr_transpose(int nrow, int ncol, int T, const int* offsets, const int* row_counts,
                                  const int* columns, int* t_offsets, int* t_row_counts, int* t_columns,
                                  int* block_map, int* status, void* workspace, size_t workspace_bytes, void* stream)
{
    return wa::bsr_transpose(nrow, ncol, T, offsets, row_counts, columns, t_offsets, t_row_counts, t_columns,
                             block_map, status, workspace, workspace_bytes, wa::as_stream(stream));
}

// Same name and arguments as the reference native entry point
// (warp/native/warp.h, wp_bsr_transpose_device; bound by name in
// warp/_src/context.py).  Runs on the stream set by
// warp_amd_set_current_stream and allocates its scratch stream-ordered.
// src_block_indices needs max(nnz, capacity) entries; errors are reported
// through warp_amd_last_error().
WA_API void wp_bsr_transpose_device(int row_count, int col_count, int nnz, const int* bsr_offsets,
                                    const int* bsr_row_counts, const int* bsr_columns, int* transposed_bsr_offsets,
                                    int* transposed_bsr_row_counts, int* transposed_bsr_columns,
                                    int* src_block_indices, int* status)
{
    hipStream_t stream = wa::as_stream(g_current_stream);
    const size_t bytes = wa::transpose_workspace_bytes(col_count, nnz);
    void* ws = nullptr;
    hipError_t err = hipMallocAsync(&ws, bytes, stream);
    if (err != hipSuccess) {
        wa::set_error("wp_bsr_transpose_device: hipMallocAsync failed: %s", hipGetErrorString(err));
        return;
    }
    wa::bsr_transpose(row_count, col_count, nnz, bsr_offsets, bsr_row_counts, bsr_columns, transposed_bsr_offsets,
                      transposed_bsr_row_counts, transposed_bsr_columns, src_block_indices, status, ws, bytes, stream);
    (void)hipFreeAsync(ws, stream);
}

// dst[b] = scale * transpose(src[block_map[b]]) for the first `count`
// transposed blocks b below *total (a device int, the transposed
// offsets[nrow]) whose block_map entry is >= 0 (reference
// _bsr_transpose_values, warp/_src/sparse.py:2604-2633).  Source blocks are
// br x bc, destination blocks bc x br; scalar_size 4 (fp32) or 8 (fp64).
WA_API int warp_amd_bsr_transpose_values(int scalar_size, int br, int bc, int64_t count, const int* total,
                                         const int* block_map, const void* src, double scale, void* dst, void* stream)
{
    WA_REQUIRE(br > 0 && bc > 0 && count >= 0, "bsr_transpose_values: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_transpose_values: values must be fp32 or fp64");
    if (count == 0)
        return wa::kOk;
    WA_REQUIRE(total && block_map && src && dst, "bsr_transpose_values: null arrays");
    hipStream_t s = wa::as_stream(stream);
    const unsigned grid = (unsigned)wa::ceil_div(count * br * bc, wa::kBlock);
    if (scalar_size == 4)
        wa::transpose_values_kernel<float><<<grid, wa::kBlock, 0, s>>>(
            count, total, br, bc, block_map, static_cast<const float*>(src), (float)scale, static_cast<float*>(dst));
    else
        wa::transpose_values_kernel<double><<<grid, wa::kBlock, 0, s>>>(
            count, total, br, bc, block_map, static_cast<const double*>(src), scale, static_cast<double*>(dst));
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Masked transpose values (reference bsr_set_transpose with
// topology="masked", warp/_src/sparse.py:2785-2792, 2831-2848): every active
// block (r, c) of the destination topology takes scale * transpose of source
// block (c, r) when the source holds one.  Destination values are not cleared
// here.
WA_API int warp_amd_bsr_transpose_masked_values(int scalar_size, int br, int bc, int t_nrow, const int* t_offsets,
                                                const int* t_row_counts, const int* t_columns, int nrow,
                                                const int* offsets, const int* row_counts, const int* columns,
                                                const void* src, double scale, void* dst, void* stream)
{
    WA_REQUIRE(br > 0 && bc > 0 && t_nrow >= 0 && nrow >= 0, "bsr_transpose_masked_values: invalid sizes");
    WA_REQUIRE(scalar_size == 4 || scalar_size == 8, "bsr_transpose_masked_values: values must be fp32 or fp64");
    if (t_nrow == 0)
        return wa::kOk;
    WA_REQUIRE(t_offsets && offsets, "bsr_transpose_masked_values: null arrays");
    hipStream_t s = wa::as_stream(stream);
    if (scalar_size == 4)
        wa::transpose_masked_values_kernel<float><<<(unsigned)t_nrow, wa::kWave, 0, s>>>(
            t_nrow, t_offsets, t_row_counts, t_columns, nrow, offsets, row_counts, columns, br, bc,
            static_cast<const float*>(src), (float)scale, static_cast<float*>(dst));
    else
        wa::transpose_masked_values_kernel<double><<<(unsigned)t_nrow, wa::kWave, 0, s>>>(
            t_nrow, t_offsets, t_row_counts, t_columns, nrow, offsets, row_counts, columns, br, bc,
            static_cast<const double*>(src), scale, static_cast<double*>(dst));
    WA_LAUNCH_CHECK();
    return wa::kOk;
}
