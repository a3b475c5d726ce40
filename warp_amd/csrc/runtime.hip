// Runtime pieces of the C ABI: error reporting, version, device queries,
// memory helpers for non-torch callers, and the int32 exclusive scan used by
// assembly.
#include "common.h"

#include <cstdarg>
#include <mutex>
#include <vector>

namespace wa {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

const char* last_error_cstr() { return g_last_error.c_str(); }

int device_cu_count()
{
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return 256;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)cache.size() <= dev)
        cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

// ---------------------------------------------------------------------------
// Exclusive scan (reduce / scan-of-partials / scan-and-add), 1024 items per
// workgroup of 256 threads.
// ---------------------------------------------------------------------------

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;
// up to this many tiles (1,048,576 items) each tile pass workgroup adds the
// raw tile sums before its own (at most 4 KB of L2 reads) instead of a third
// launch scanning them
constexpr int64_t kScanRawTiles = 1024;

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const int* in, int64_t n, int* partials)
{
    __shared__ int lds[kScanBlock / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
        if (i < n)
            s += in[i];
    }
    s = block_sum<kScanBlock>(s, lds);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// Exclusive scan of `count` partials in place by a single workgroup; writes
// the grand total to *total.
__global__ __launch_bounds__(1024) void scan_partials_kernel(int* partials, int64_t count, int* total)
{
    __shared__ int lds[1024];
    int carry = 0;
    for (int64_t base = 0; base < count; base += 1024) {
        int64_t i = base + threadIdx.x;
        int v = i < count ? partials[i] : 0;
        lds[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int t = threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
            __syncthreads();
            lds[threadIdx.x] += t;
            __syncthreads();
        }
        int incl = lds[threadIdx.x];
        if (i < count)
            partials[i] = carry + incl - v;
        int tile_total = lds[1023];
        __syncthreads();
        carry += tile_total;
    }
    if (threadIdx.x == 0)
        *total = carry;
}

// RAW: `partials` hold the tiles' plain sums; each workgroup adds those of
// the tiles before it (few tiles: this replaces the one-workgroup scan of the
// partials) and the last one writes the total to out[n].  Otherwise
// `partials` hold the scanned tile offsets.
// CLEAR: zero `in` once the tile is loaded (in != out).
template <bool RAW, bool CLEAR>
__global__ __launch_bounds__(kScanBlock) void scan_tile_kernel(
    int* in, int* out, int64_t n, const int* partials
)
{
    __shared__ int lds[kScanTile];
    __shared__ int wsum[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int prefix = 0;
    if constexpr (RAW) {
        int a = 0;
        for (int k = threadIdx.x; k < (int)blockIdx.x; k += kScanBlock)
            a += partials[k];
        prefix = block_sum<kScanBlock>(a, wsum);
    } else {
        prefix = partials[blockIdx.x];
    }
    // load the tile (coalesced), each thread then owns kScanItems consecutive items
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
        lds[k * kScanBlock + threadIdx.x] = i < n ? in[i] : 0;
        if (CLEAR && i < n)
            in[i] = 0;
    }
    __syncthreads();
    int local[kScanItems];
    int run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        local[k] = run;
        run += lds[threadIdx.x * kScanItems + k];
    }
    wsum[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int t = threadIdx.x >= off ? wsum[threadIdx.x - off] : 0;
        __syncthreads();
        wsum[threadIdx.x] += t;
        __syncthreads();
    }
    const int excl = wsum[threadIdx.x] - run + prefix;
    if (RAW && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1)
        out[n] = excl + run;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        lds[threadIdx.x * kScanItems + k] = excl + local[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
        if (i < n)
            out[i] = lds[k * kScanBlock + threadIdx.x];
    }
}

__global__ void scan_empty_kernel(int* out) { *out = 0; }

__global__ __launch_bounds__(256) void fill_i32_kernel(int* p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = v;
}

}  // namespace

int launch_fill_i32(int* p, int64_t n, int v, hipStream_t stream)
{
    if (n <= 0)
        return kOk;
    fill_i32_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, stream>>>(p, n, v);
    WA_LAUNCH_CHECK();
    return kOk;
}

size_t scan_workspace_bytes(int64_t n)
{
    int64_t tiles = ceil_div(n > 0 ? n : 1, kScanTile);
    return align_up((size_t)tiles * sizeof(int), 256);
}

namespace {

template <bool CLEAR>
int scan_i32(int* in, int* out, int64_t n, void* ws, hipStream_t stream)
{
    if (n <= 0) {
        // a kernel rather than a memset node, so the call can be graph-captured
        scan_empty_kernel<<<1, 1, 0, stream>>>(out);
        WA_LAUNCH_CHECK();
        return kOk;
    }
    const int64_t tiles = ceil_div(n, kScanTile);
    int* partials = static_cast<int*>(ws);
    // The reduce pass reads `in` before the tile pass overwrites it, so in-place use is safe.
    scan_reduce_kernel<<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, n, partials);
    WA_LAUNCH_CHECK();
    if (tiles <= kScanRawTiles) {
        scan_tile_kernel<true, CLEAR><<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, out, n, partials);
        WA_LAUNCH_CHECK();
        return kOk;
    }
    scan_partials_kernel<<<1, 1024, 0, stream>>>(partials, tiles, out + n);
    WA_LAUNCH_CHECK();
    scan_tile_kernel<false, CLEAR><<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, out, n, partials);
    WA_LAUNCH_CHECK();
    return kOk;
}

}  // namespace

int exclusive_scan_i32(const int* in, int* out, int64_t n, void* ws, hipStream_t stream)
{
    // read-only use of `in` (CLEAR is off), which may alias `out`
    return scan_i32<false>(const_cast<int*>(in), out, n, ws, stream);
}

int exclusive_scan_i32_clear(int* in, int* out, int64_t n, void* ws, hipStream_t stream)
{
    WA_REQUIRE(in != out, "exclusive_scan_i32_clear: input and output must differ");
    return scan_i32<true>(in, out, n, ws, stream);
}

}  // namespace wa

// ---------------------------------------------------------------------------
// C ABI: runtime
// ---------------------------------------------------------------------------

WA_API const char* warp_amd_last_error(void) { return wa::last_error_cstr(); }

WA_API void warp_amd_clear_error(void) { wa::clear_error(); }

// Library version: major * 10000 + minor * 100 + patch.
WA_API int warp_amd_version(void) { return 100; }

WA_API int warp_amd_device_count(int* count)
{
    WA_HIP(hipGetDeviceCount(count));
    return wa::kOk;
}

WA_API int warp_amd_malloc(void** ptr, size_t bytes)
{
    WA_HIP(hipMalloc(ptr, bytes));
    return wa::kOk;
}

WA_API int warp_amd_free(void* ptr)
{
    WA_HIP(hipFree(ptr));
    return wa::kOk;
}

WA_API int warp_amd_malloc_host(void** ptr, size_t bytes)
{
    WA_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return wa::kOk;
}

WA_API int warp_amd_free_host(void* ptr)
{
    WA_HIP(hipHostFree(ptr));
    return wa::kOk;
}

// kind: 0 host->device, 1 device->host, 2 device->device
WA_API int warp_amd_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream)
{
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    WA_REQUIRE(kind >= 0 && kind < 3, "warp_amd_memcpy: invalid kind %d", kind);
    WA_HIP(hipMemcpyAsync(dst, src, bytes, kinds[kind], wa::as_stream(stream)));
    return wa::kOk;
}

WA_API int warp_amd_stream_synchronize(void* stream)
{
    WA_HIP(hipStreamSynchronize(wa::as_stream(stream)));
    return wa::kOk;
}
