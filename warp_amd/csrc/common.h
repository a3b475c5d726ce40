// Shared definitions for the gfx950 sparse-solver library.
//
// Conventions of the C ABI (include/warp_amd.h): every entry point returns an
// int status (0 = success, < 0 = HIP or argument error) and records a message
// readable through warp_amd_last_error().  Work is enqueued on the stream
// passed as `void* stream` (null = the legacy default stream) and is never
// synchronised unless the entry point documents a host read-back.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#define WA_API extern "C" __attribute__((visibility("default")))

namespace wa {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

enum Status : int {
    kOk = 0,
    kErrHip = -1,
    kErrArg = -2,
    kErrUnsupported = -3,
    kErrWorkspace = -4,
};

// Sticky device status codes of padded-topology operations (reference
// BSR_STATUS_*, warp/_src/sparse.py:80-84).
constexpr int kStatusSuccess = 0;
constexpr int kStatusRowCapacityExceeded = 1;

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
const char* last_error_cstr();

#define WA_HIP(call)                                                                                                  \
    do {                                                                                                              \
        hipError_t wa_err_ = (call);                                                                                  \
        if (wa_err_ != hipSuccess) {                                                                                  \
            ::wa::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(wa_err_));           \
            return ::wa::kErrHip;                                                                                     \
        }                                                                                                             \
    } while (0)

#define WA_LAUNCH_CHECK()                                                                                             \
    do {                                                                                                              \
        hipError_t wa_err_ = hipGetLastError();                                                                       \
        if (wa_err_ != hipSuccess) {                                                                                  \
            ::wa::set_error("%s:%d: kernel launch failed: %s", __FILE__, __LINE__, hipGetErrorString(wa_err_));       \
            return ::wa::kErrHip;                                                                                     \
        }                                                                                                             \
    } while (0)

#define WA_REQUIRE(cond, ...)                                                                                         \
    do {                                                                                                              \
        if (!(cond)) {                                                                                                \
            ::wa::set_error(__VA_ARGS__);                                                                             \
            return ::wa::kErrArg;                                                                                     \
        }                                                                                                             \
    } while (0)

#define WA_TRY(expr)                                                                                                  \
    do {                                                                                                              \
        int wa_st_ = (expr);                                                                                          \
        if (wa_st_ != 0)                                                                                              \
            return wa_st_;                                                                                            \
    } while (0)

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace (256-byte aligned slices).
struct Carve {
    char* base;
    size_t used = 0;
    size_t cap;
    Carve(void* p, size_t c) : base(static_cast<char*>(p)), cap(c) {}
    template <typename T> T* take(size_t count)
    {
        size_t off = align_up(used, 256);
        used = off + count * sizeof(T);
        return base ? reinterpret_cast<T*>(base + off) : nullptr;
    }
    bool fits() const { return used <= cap; }
};

// Number of compute units of the current device (cached per device id).
int device_cu_count();

// Lanes per block row used by the vector-BSR kernels (sparse_mv.hip).
int bsr_mv_lanes(int nrow, int br, int bc, int64_t nnz_hint);

// Exclusive scan of n int32 values: out[i] = sum(in[0..i)), out[n] = total.
// `in` and `out` may alias.  Workspace from scan_workspace_bytes(n).
size_t scan_workspace_bytes(int64_t n);
// p[0..n) = v on `stream` (a kernel, never a memset node)
int launch_fill_i32(int* p, int64_t n, int v, hipStream_t stream);
int exclusive_scan_i32(const int* in, int* out, int64_t n, void* ws, hipStream_t stream);
// The same with `in` (distinct from `out`) left zeroed: out = scan(in), in = 0.
int exclusive_scan_i32_clear(int* in, int* out, int64_t n, void* ws, hipStream_t stream);

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over aligned groups of G lanes (G a power of two <= 64).
template <int G, typename T> __device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum of one value per thread; result valid in every thread.
template <int BLOCK, typename T> __device__ __forceinline__ T block_sum(T v, T* lds /* >= BLOCK/64 entries */)
{
    v = wave_sum(v);
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (lane == 0)
        lds[wave] = v;
    __syncthreads();
    T total = T(0);
#pragma unroll
    for (int w = 0; w < BLOCK / kWave; ++w)
        total += lds[w];
    __syncthreads();
    return total;
}

// Device-coherent loads/stores for values exchanged between workgroups of
// one kernel (bypass the per-CU vector cache).
template <typename T> __device__ __forceinline__ T load_agent(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void store_agent(T* p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block partial sums for a following finalize kernel: thread 0 of workgroup b
// writes the NV workgroup sums to partials[v * gridDim.x + b] with plain
// stores (the kernel boundary makes them visible; no per-workgroup fences,
// which on a multi-XCD GPU flush the L2 and serialise the grid).
template <int BLOCK, int NV>
__device__ __forceinline__ void write_block_partials(const double (&mine)[NV], double* partials, double* lds)
{
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double s = block_sum<BLOCK>(mine[v], lds);
        if (threadIdx.x == 0)
            partials[(size_t)v * gridDim.x + blockIdx.x] = s;
    }
}

// Fixed-order sums of NV rows of `count` partials (row v at partials + v *
// stride) by one workgroup; results in every thread.  Each thread adds its
// entries k = tid, tid + BLOCK, ... in ascending order, but loads them in
// batches of kBatch per row first (64 VGPRs of loads in flight): a finalize kernel is one workgroup, so its
// time is the number of dependent L2 round trips, not bandwidth.
template <int BLOCK, int NV>
__device__ __forceinline__ void sum_partials_rows(const double* partials, size_t stride, int count, double* lds,
                                                  double (&out)[NV])
{
    constexpr int kBatch = 32 / NV > 4 ? 32 / NV : 4;
    double a[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v)
        a[v] = 0.0;
    int k = threadIdx.x;
    for (; k + (kBatch - 1) * BLOCK < count; k += kBatch * BLOCK) {
        double t[NV][kBatch];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                t[v][u] = partials[(size_t)v * stride + k + u * BLOCK];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                a[v] += t[v][u];
    }
    for (; k < count; k += BLOCK)
#pragma unroll
        for (int v = 0; v < NV; ++v)
            a[v] += partials[(size_t)v * stride + k];
#pragma unroll
    for (int v = 0; v < NV; ++v)
        out[v] = block_sum<BLOCK>(a[v], lds);
}

// Fixed-order sum of `count` partials by one workgroup (result in all threads).
template <int BLOCK>
__device__ __forceinline__ double sum_partials(const double* partials, int count, double* lds)
{
    double out[1];
    sum_partials_rows<BLOCK, 1>(partials, 0, count, lds, out);
    return out[0];
}

}  // namespace wa
