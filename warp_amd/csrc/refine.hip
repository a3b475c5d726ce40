// Mixed-precision pieces for FP64 iterative refinement around the FP32 CG,
// and the vector kernels of the flexible CG with the learned preconditioner.
//
// No reference counterpart: the reference solves in one precision
// (warp/_src/optim/linear.py).  Classical iterative refinement (e.g. Carson &
// Higham, SIAM J. Sci. Comput. 40(2), 2018): r = b - A x in FP64 with the FP32
// matrix promoted in-kernel, solve A d = r in FP32, x += d in FP64.
#include "bsr_core.h"

#include <mutex>

namespace wa {
namespace {

constexpr int kBlock = 256;
constexpr int kFinalizeBlock = 1024;

// With b: r64 = b - A x (FP64 accumulate, FP32 values), r32 = (float) r64
// (when non-null), partials of r.r.  Without b: r64 = A x, partials of x.(A x).
template <int BR, int G, int RPG>
__global__ __launch_bounds__(kBlock) void residual_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ columns, const float* __restrict__ values,
    const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r64, float* __restrict__ r32,
    double* __restrict__ partials
)
{
    __shared__ double lds[kBlock / kWave];
    const int row0 = (blockIdx.x * (kBlock / G) + threadIdx.x / G) * RPG;
    const int lane = threadIdx.x % G;
    double acc[1] = {0.0};
    if (row0 < nrow) {
        const RowAcc<double, BR * RPG> ax =
            group_rows_product<double, BR, BR, G, RPG, float>(row0, nrow, lane, offsets, nullptr, columns, values, x);
        for (int e = lane; e < BR * RPG && row0 + e / BR < nrow; e += G) {
            const size_t i = (size_t)row0 * BR + e;
            if (b) {
                const double rv = b[i] - pick(ax, e);
                r64[i] = rv;
                if (r32)
                    r32[i] = (float)rv;
                acc[0] += rv * rv;
            } else {
                const double yv = pick(ax, e);
                r64[i] = yv;
                acc[0] += x[i] * yv;
            }
        }
    }
    write_block_partials<kBlock, 1>(acc, partials, lds);
}

// One damped Jacobi sweep out of place: z_out = z + omega D^-1 (r - A z)
// (FP32 matrix promoted, FP64 vectors).
template <int BR, int G, int RPG>
__global__ __launch_bounds__(kBlock) void jacobi_sweep_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ columns, const float* __restrict__ values,
    const double* __restrict__ z, const double* __restrict__ r, const double* __restrict__ dinv, double omega,
    double* __restrict__ z_out
)
{
    const int row0 = (blockIdx.x * (kBlock / G) + threadIdx.x / G) * RPG;
    const int lane = threadIdx.x % G;
    if (row0 >= nrow)
        return;
    const RowAcc<double, BR * RPG> az =
        group_rows_product<double, BR, BR, G, RPG, float>(row0, nrow, lane, offsets, nullptr, columns, values, z);
    for (int e = lane; e < BR * RPG && row0 + e / BR < nrow; e += G) {
        const size_t i = (size_t)row0 * BR + e;
        z_out[i] = fma(omega * dinv[i], r[i] - pick(az, e), z[i]);
    }
}

// ---------------------------------------------------------------------------
// Fused flexible-CG iteration for a scalar (1x1) FP32 matrix, FP64 vectors,
// one Jacobi sweep before and after the network correction.  Vectors the
// SpMV gathers are formed on the fly from their parts (XF functors), so the
// sweep results are never written and re-read.
// ---------------------------------------------------------------------------

struct WeightX {  // w = omega D^-1 r, kept in FP32 by the update kernel
    const float* __restrict__ w;
    __device__ __forceinline__ double operator()(size_t i) const { return (double)w[i]; }
};

struct CorrectedX {  // z2 = w + c * out
    const float* __restrict__ w;
    const float* __restrict__ out;
    double c;
    __device__ __forceinline__ double operator()(size_t i) const { return fma(c, (double)out[i], (double)w[i]); }
};

struct DirectionX {  // z - beta p_prev
    const float* __restrict__ z;
    const double* __restrict__ p_prev;
    double beta;
    __device__ __forceinline__ double operator()(size_t i) const { return fma(-beta, p_prev[i], (double)z[i]); }
};

// r1 = r - A w; partials of r1.r1
template <int G, int RPG>
__global__ __launch_bounds__(kBlock) void fcg_pre_kernel(int nrow, const int* __restrict__ offsets,
                                                        const int* __restrict__ columns,
                                                        const float* __restrict__ values, const double* __restrict__ r,
                                                        const float* __restrict__ w, double* __restrict__ r1,
                                                        double* __restrict__ partials)
{
    __shared__ double lds[kBlock / kWave];
    const int row0 = (blockIdx.x * (kBlock / G) + threadIdx.x / G) * RPG;
    const int lane = threadIdx.x % G;
    double acc[1] = {0.0};
    if (row0 < nrow) {
        const RowAcc<double, RPG> aw = group_rows_product_f<double, 1, 1, G, RPG, float>(
            row0, nrow, lane, offsets, nullptr, columns, values, WeightX{w}
        );
        for (int e = lane; e < RPG && row0 + e < nrow; e += G) {
            const size_t i = (size_t)row0 + e;
            const double v = r[i] - pick(aw, e);
            r1[i] = v;
            acc[0] += v * v;
        }
    }
    write_block_partials<kBlock, 1>(acc, partials, lds);
}

// z2 = w + gain * scale * out;  z = z2 + omega D^-1 (r - A z2) (stored FP32);
// partials of z.q_prev
template <int G, int RPG>
__global__ __launch_bounds__(kBlock) void fcg_post_kernel(int nrow, const int* __restrict__ offsets,
                                                         const int* __restrict__ columns,
                                                         const float* __restrict__ values, const double* __restrict__ r,
                                                         const double* __restrict__ dinv, double omega,
                                                         const float* __restrict__ w, const float* __restrict__ out,
                                                         const double* scale, double gain,
                                                         const double* __restrict__ q_prev, float* __restrict__ z,
                                                         double* __restrict__ partials)
{
    __shared__ double lds[kBlock / kWave];
    const int row0 = (blockIdx.x * (kBlock / G) + threadIdx.x / G) * RPG;
    const int lane = threadIdx.x % G;
    double acc[1] = {0.0};
    if (row0 < nrow) {
        const CorrectedX xf{w, out, gain * (*scale)};
        const RowAcc<double, RPG> az =
            group_rows_product_f<double, 1, 1, G, RPG, float>(row0, nrow, lane, offsets, nullptr, columns, values, xf);
        for (int e = lane; e < RPG && row0 + e < nrow; e += G) {
            const size_t i = (size_t)row0 + e;
            const float v = (float)fma(omega * dinv[i], r[i] - pick(az, e), xf(i));
            z[i] = v;
            acc[0] += (double)v * q_prev[i];
        }
    }
    write_block_partials<kBlock, 1>(acc, partials, lds);
}

// beta = zq / pq_prev;  p = z - beta p_prev;  q = A p;  partials of p.q, p.r
template <int G, int RPG>
__global__ __launch_bounds__(kBlock) void fcg_dir_spmv_kernel(int nrow, const int* __restrict__ offsets,
                                                             const int* __restrict__ columns,
                                                             const float* __restrict__ values,
                                                             const float* __restrict__ z,
                                                             const double* __restrict__ p_prev, const double* zq,
                                                             const double* pq_prev, const double* __restrict__ r,
                                                             double* __restrict__ p, double* __restrict__ q,
                                                             double* __restrict__ partials)
{
    __shared__ double lds[kBlock / kWave];
    const int row0 = (blockIdx.x * (kBlock / G) + threadIdx.x / G) * RPG;
    const int lane = threadIdx.x % G;
    double acc[2] = {0.0, 0.0};
    if (row0 < nrow) {
        const double beta = *pq_prev != 0.0 ? *zq / *pq_prev : 0.0;
        const DirectionX xf{z, p_prev, beta};
        const RowAcc<double, RPG> ap =
            group_rows_product_f<double, 1, 1, G, RPG, float>(row0, nrow, lane, offsets, nullptr, columns, values, xf);
        for (int e = lane; e < RPG && row0 + e < nrow; e += G) {
            const size_t i = (size_t)row0 + e;
            const double pv = xf(i);
            const double qv = pick(ap, e);
            p[i] = pv;
            q[i] = qv;
            acc[0] += pv * qv;
            acc[1] += pv * r[i];
        }
    }
    write_block_partials<kBlock, 2>(acc, partials, lds);
}

// w = (float)(omega * dinv * r)
__global__ __launch_bounds__(kBlock) void fcg_weight_kernel(int64_t n, const double* __restrict__ dinv,
                                                           const double* __restrict__ r, double omega,
                                                           float* __restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        w[i] = (float)(omega * dinv[i] * r[i]);
}

// out0 = sum of partials[0, nparts); with out1: out1 = sum of partials[nparts, 2 nparts);
// with scale/inv_scale: scale = sqrt(out0 / n), inv_scale = 1 / scale (0 if scale is 0)
__global__ __launch_bounds__(kFinalizeBlock) void finalize_fcg_kernel(const double* partials, int nparts,
                                                                     double* out0, double* out1, int64_t n,
                                                                     double* scale, double* inv_scale)
{
    __shared__ double lds[kFinalizeBlock / kWave];
    double t0, t1 = 0.0;
    if (out1) {
        double t[2];
        sum_partials_rows<kFinalizeBlock, 2>(partials, (size_t)nparts, nparts, lds, t);
        t0 = t[0];
        t1 = t[1];
    } else {
        t0 = sum_partials<kFinalizeBlock>(partials, nparts, lds);
    }
    if (threadIdx.x == 0) {
        *out0 = t0;
        if (out1)
            *out1 = t1;
        if (scale) {
            const double sv = sqrt(t0 / (double)n);
            *scale = sv;
            *inv_scale = sv > 0.0 ? 1.0 / sv : 0.0;
        }
    }
}

// z += mult * (*scale) * out
__global__ __launch_bounds__(kBlock) void unpack_add_kernel(int64_t n, const float* __restrict__ out, const double* scale,
                                                           double mult, double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        z[i] = fma(mult * (*scale), (double)out[i], z[i]);
}

// One-step energy reduction of a search direction z for the residual r of
// A e = r:  L = 1 - (r.z)^2 / ((r.e)(z.A z)), the fraction of the A-norm error
// left after the exact line search along z.  With sc = {r.z, z.Az, r.e}:
// g = dL/dz = -2 a / (c b) r + 2 a^2 / (c b^2) A z, and loss = L.
__global__ __launch_bounds__(kBlock) void energy_grad_kernel(int64_t n, const double* __restrict__ r,
                                                            const double* __restrict__ az, const double* sc,
                                                            double* __restrict__ g, double* loss)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double a = sc[0], b = sc[1], c = sc[2];
    const double cb = c * b;
    const double ca = cb != 0.0 ? -2.0 * a / cb : 0.0;
    const double cz = (cb != 0.0 && b != 0.0) ? 2.0 * a * a / (cb * b) : 0.0;
    if (i == 0)
        *loss = cb != 0.0 ? 1.0 - a * a / cb : 1.0;
    if (i < n)
        g[i] = fma(ca, r[i], cz * az[i]);
}

// out32 = (float)(mult * (*scale) * x)
__global__ __launch_bounds__(kBlock) void scale_to_f32_kernel(int64_t n, const double* __restrict__ x, const double* scale,
                                                             double mult, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        out[i] = (float)(mult * (*scale) * x[i]);
}

// scale = sqrt(ss / n), inv_scale = 1 / scale (0 when scale is 0)
__global__ void rms_scales_kernel(const double* ss, int64_t n, double* scale, double* inv_scale)
{
    const double s = sqrt(*ss / (double)n);
    *scale = s;
    *inv_scale = s > 0.0 ? 1.0 / s : 0.0;
}

// Flexible CG direction (Notay 2000, truncation 1): p = z - (zq / pq_prev) p_prev
__global__ __launch_bounds__(kBlock) void fcg_direction_kernel(
    int64_t n, const double* __restrict__ z, const double* __restrict__ p_prev, const double* zq,
    const double* pq_prev, double* __restrict__ p
)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const double beta = *pq_prev != 0.0 ? *zq / *pq_prev : 0.0;
    p[i] = fma(-beta, p_prev[i], z[i]);
}

// alpha = pr / pq; x += alpha p; r -= alpha q; partials of r.r
// (with w: also w = (float)(omega * dinv * r), the next pre-smoothing sweep)
__global__ __launch_bounds__(kBlock) void fcg_update_kernel(
    int64_t n, const double* pr, const double* pq, const double* __restrict__ p, const double* __restrict__ q,
    double* __restrict__ x, double* __restrict__ r, const double* __restrict__ dinv, double omega,
    float* __restrict__ w, double* __restrict__ partials
)
{
    __shared__ double lds[kBlock / kWave];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double alpha = *pq != 0.0 ? *pr / *pq : 0.0;
    double acc[1] = {0.0};
    if (i < n) {
        x[i] = fma(alpha, p[i], x[i]);
        const double rv = fma(-alpha, q[i], r[i]);
        r[i] = rv;
        if (w)
            w[i] = (float)(omega * dinv[i] * rv);
        acc[0] = rv * rv;
    }
    write_block_partials<kBlock, 1>(acc, partials, lds);
}

__global__ __launch_bounds__(kFinalizeBlock) void finalize_sum_kernel(const double* partials, int nparts, double* out)
{
    __shared__ double lds[kFinalizeBlock / kWave];
    const double t = sum_partials<kFinalizeBlock>(partials, nparts, lds);
    if (threadIdx.x == 0)
        *out = t;
}

__global__ __launch_bounds__(kBlock) void axpy_f32_to_f64_kernel(int64_t n, double alpha, const float* __restrict__ d,
                                                                 double* __restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        x[i] = fma(alpha, (double)d[i], x[i]);
}

template <int BR, int G>
int launch_residual(
    int nrow, const int* offsets, const int* columns, const float* values, const double* x, const double* b,
    double* r64, float* r32, double* partials, double* rr, hipStream_t stream
)
{
    constexpr int RPG = rows_per_group_for(BR, BR);
    const unsigned grid = (unsigned)ceil_div(nrow, (kBlock / G) * RPG);
    residual_kernel<BR, G, RPG><<<grid, kBlock, 0, stream>>>(nrow, offsets, columns, values, x, b, r64, r32, partials);
    WA_LAUNCH_CHECK();
    finalize_sum_kernel<<<1, kFinalizeBlock, 0, stream>>>(partials, (int)grid, rr);
    WA_LAUNCH_CHECK();
    return kOk;
}

template <int BR>
int residual_dispatch(
    int g, int nrow, const int* offsets, const int* columns, const float* values, const double* x, const double* b,
    double* r64, float* r32, double* partials, double* rr, hipStream_t stream
)
{
    switch (g) {
    case 1: return launch_residual<BR, 1>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    case 2: return launch_residual<BR, 2>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    case 4: return launch_residual<BR, 4>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    case 8: return launch_residual<BR, 8>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    case 16: return launch_residual<BR, 16>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    case 32: return launch_residual<BR, 32>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    default: return launch_residual<BR, 64>(nrow, offsets, columns, values, x, b, r64, r32, partials, rr, stream);
    }
}

template <int G>
struct FcgLaunch {
    static constexpr int RPG = rows_per_group_for(1, 1);
    static unsigned grid(int nrow) { return (unsigned)ceil_div(nrow, (kBlock / G) * RPG); }
    static int pre(int nrow, const int* offsets, const int* columns, const float* values, const double* r,
                   const float* w, double* r1, double* rr1, double* scale, double* inv_scale, double* partials,
                   hipStream_t stream)
    {
        const unsigned g = grid(nrow);
        fcg_pre_kernel<G, RPG><<<g, kBlock, 0, stream>>>(nrow, offsets, columns, values, r, w, r1, partials);
        WA_LAUNCH_CHECK();
        finalize_fcg_kernel<<<1, kFinalizeBlock, 0, stream>>>(partials, (int)g, rr1, nullptr, nrow, scale, inv_scale);
        WA_LAUNCH_CHECK();
        return kOk;
    }
    static int post(int nrow, const int* offsets, const int* columns, const float* values, const double* r,
                    const double* dinv, double omega, const float* w, const float* out, const double* scale,
                    double gain, const double* q_prev, float* z, double* zq, double* partials, hipStream_t stream)
    {
        const unsigned g = grid(nrow);
        fcg_post_kernel<G, RPG><<<g, kBlock, 0, stream>>>(nrow, offsets, columns, values, r, dinv, omega, w, out,
                                                           scale, gain, q_prev, z, partials);
        WA_LAUNCH_CHECK();
        finalize_fcg_kernel<<<1, kFinalizeBlock, 0, stream>>>(partials, (int)g, zq, nullptr, nrow, nullptr, nullptr);
        WA_LAUNCH_CHECK();
        return kOk;
    }
    static int dir(int nrow, const int* offsets, const int* columns, const float* values, const float* z,
                   const double* p_prev, const double* zq, const double* pq_prev, const double* r, double* p, double* q,
                   double* pq, double* pr, double* partials, hipStream_t stream)
    {
        const unsigned g = grid(nrow);
        fcg_dir_spmv_kernel<G, RPG><<<g, kBlock, 0, stream>>>(nrow, offsets, columns, values, z, p_prev, zq, pq_prev,
                                                               r, p, q, partials);
        WA_LAUNCH_CHECK();
        finalize_fcg_kernel<<<1, kFinalizeBlock, 0, stream>>>(partials, (int)g, pq, pr, nrow, nullptr, nullptr);
        WA_LAUNCH_CHECK();
        return kOk;
    }
};

template <int G>
int launch_sweep(int nrow, const int* offsets, const int* columns, const float* values, const double* z,
                 const double* r, const double* dinv, double omega, double* z_out, hipStream_t stream)
{
    constexpr int RPG = rows_per_group_for(1, 1);
    const unsigned grid = (unsigned)ceil_div(nrow, (kBlock / G) * RPG);
    jacobi_sweep_kernel<1, G, RPG><<<grid, kBlock, 0, stream>>>(nrow, offsets, columns, values, z, r, dinv, omega, z_out);
    WA_LAUNCH_CHECK();
    return kOk;
}

}  // namespace
}  // namespace wa

// Partials of the residual kernel: one double per workgroup, at most one
// workgroup per 4 block rows.
WA_API size_t warp_amd_bsr_residual_workspace_size(int nrow)
{
    return sizeof(double) * (size_t)(wa::ceil_div(nrow > 0 ? nrow : 1, 4) + 1);
}

// r64 = b - A x and r32 = (float) r64 for a compact BSR matrix with FP32
// values (1x1 or 3x3 blocks) and FP64 vectors; *rr (device) = r64 . r64,
// summed in a fixed order.
WA_API int warp_amd_bsr_residual_f32_f64(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    const double* x, const double* b, double* r64, float* r32, double* rr, void* workspace, size_t workspace_bytes,
    void* stream
)
{
    WA_REQUIRE(block_size == 1 || block_size == 3, "bsr_residual: block size must be 1 or 3, got %d", block_size);
    WA_REQUIRE(nrow > 0, "bsr_residual: empty matrix");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_bsr_residual_workspace_size(nrow),
               "bsr_residual: workspace too small");
    const int g = wa::default_bsr_shape(nrow, block_size, block_size, nnz_hint).lanes;
    double* partials = static_cast<double*>(workspace);
    if (block_size == 1)
        return wa::residual_dispatch<1>(g, nrow, offsets, columns, values, x, b, r64, r32, partials, rr,
                                        wa::as_stream(stream));
    return wa::residual_dispatch<3>(g, nrow, offsets, columns, values, x, b, r64, r32, partials, rr,
                                    wa::as_stream(stream));
}

// ---------------------------------------------------------------------------
// The residual without a workspace, for callers that cannot allocate (code
// under graph capture, serving loops): its partials live in `scratch`, a
// file-scope device array that every GPU the library runs on holds until the
// process ends.  Calls on one GPU share it, so calls that overlap on one GPU
// race.  Every partial a call sums is written by that call's residual
// kernel, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------

namespace wa {

// 65,536 FP64 partials (512 KB): one per workgroup of the residual kernel,
// i.e. up to 16,777,216 scalar rows or 1,048,576 block rows of 3x3 blocks at
// the default lanes per row.
constexpr int kResidualScratchParts = 65536;

__device__ __align__(256) double scratch[kResidualScratchParts];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy at its own address; the host-side
// name is not a device address).
int scratch_on_current_device(double** out)
{
    int dev = -1;
    WA_HIP(hipGetDevice(&dev));
    WA_REQUIRE(dev >= 0 && dev < kMaxDevices, "bsr_residual_noalloc: device id %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        WA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return kOk;
}

}  // namespace
}  // namespace wa

// warp_amd_bsr_residual_f32_f64 with the partials in the library's device
// scratch: allocates nothing.  Sizes whose residual grid exceeds the scratch
// are rejected before any launch.
WA_API int warp_amd_bsr_residual_f32_f64_noalloc(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    const double* x, const double* b, double* r64, float* r32, double* rr, void* stream
)
{
    WA_REQUIRE(block_size == 1 || block_size == 3, "bsr_residual_noalloc: block size must be 1 or 3, got %d",
               block_size);
    WA_REQUIRE(nrow > 0, "bsr_residual_noalloc: empty matrix");
    const wa::BsrShape shape = wa::default_bsr_shape(nrow, block_size, block_size, nnz_hint);
    const int64_t rows_per_block = (int64_t)(wa::kBlock / shape.lanes) * wa::rows_per_group_for(block_size, block_size);
    const int64_t parts = wa::ceil_div(nrow, rows_per_block);
    WA_REQUIRE(parts <= wa::kResidualScratchParts,
               "bsr_residual_noalloc: %d block rows need %lld partials, more than the %d the scratch holds; "
               "use warp_amd_bsr_residual_f32_f64 with a workspace",
               nrow, (long long)parts, wa::kResidualScratchParts);
    double* partials = nullptr;
    WA_TRY(wa::scratch_on_current_device(&partials));
    if (block_size == 1)
        return wa::residual_dispatch<1>(shape.lanes, nrow, offsets, columns, values, x, b, r64, r32, partials, rr,
                                        wa::as_stream(stream));
    return wa::residual_dispatch<3>(shape.lanes, nrow, offsets, columns, values, x, b, r64, r32, partials, rr,
                                    wa::as_stream(stream));
}

// y64 = A x64 (FP32 values promoted, FP64 accumulation) and *xy (device) = x.y
WA_API int warp_amd_bsr_mv_f32_f64(
    int nrow, int block_size, const int* offsets, const int* columns, const float* values, int64_t nnz_hint,
    const double* x, double* y, double* xy, void* workspace, size_t workspace_bytes, void* stream
)
{
    WA_REQUIRE(block_size == 1 || block_size == 3, "bsr_mv_f32_f64: block size must be 1 or 3, got %d", block_size);
    WA_REQUIRE(nrow > 0, "bsr_mv_f32_f64: empty matrix");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_bsr_residual_workspace_size(nrow),
               "bsr_mv_f32_f64: workspace too small");
    const int g = wa::default_bsr_shape(nrow, block_size, block_size, nnz_hint).lanes;
    double* partials = static_cast<double*>(workspace);
    if (block_size == 1)
        return wa::residual_dispatch<1>(g, nrow, offsets, columns, values, x, nullptr, y, nullptr, partials, xy,
                                        wa::as_stream(stream));
    return wa::residual_dispatch<3>(g, nrow, offsets, columns, values, x, nullptr, y, nullptr, partials, xy,
                                    wa::as_stream(stream));
}

WA_API int warp_amd_rms_scales(const double* ss, int64_t n, double* scale, double* inv_scale, void* stream)
{
    wa::rms_scales_kernel<<<1, 1, 0, wa::as_stream(stream)>>>(ss, n, scale, inv_scale);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_fcg_direction(int64_t n, const double* z, const double* p_prev, const double* zq,
                                  const double* pq_prev, double* p, void* stream)
{
    wa::fcg_direction_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, z, p_prev, zq, pq_prev, p
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Workspace: one double per 256 entries.
WA_API size_t warp_amd_fcg_update_workspace_size(int64_t n) { return sizeof(double) * (size_t)(wa::ceil_div(n, 256) + 1); }

WA_API int warp_amd_fcg_update(int64_t n, const double* pr, const double* pq, const double* p, const double* q,
                               double* x, double* r, double* rr, const double* dinv, double omega, float* w,
                               void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(!w || dinv, "fcg_update: the weighted residual needs dinv");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_fcg_update_workspace_size(n), "fcg_update: workspace too small");
    hipStream_t s = wa::as_stream(stream);
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    double* partials = static_cast<double*>(workspace);
    wa::fcg_update_kernel<<<grid, wa::kBlock, 0, s>>>(n, pr, pq, p, q, x, r, dinv, omega, w, partials);
    WA_LAUNCH_CHECK();
    wa::finalize_sum_kernel<<<1, wa::kFinalizeBlock, 0, s>>>(partials, (int)grid, rr);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// x64 += alpha * (double) d32
WA_API int warp_amd_axpy_f32_to_f64(int64_t n, double alpha, const float* d, double* x, void* stream)
{
    if (n <= 0)
        return wa::kOk;
    wa::axpy_f32_to_f64_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, alpha, d, x
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// One damped Jacobi sweep for a scalar (1x1) compact BSR matrix with FP32
// values: z_out = z + omega * dinv * (r - A z); z_out must not alias z.
WA_API int warp_amd_jacobi_sweep_f32_f64(int nrow, const int* offsets, const int* columns, const float* values,
                                         int64_t nnz_hint, const double* z, const double* r, const double* dinv,
                                         double omega, double* z_out, void* stream)
{
    WA_REQUIRE(nrow > 0, "jacobi_sweep: empty matrix");
    WA_REQUIRE(z_out != z, "jacobi_sweep: z_out must not alias z");
    const int g = wa::default_bsr_shape(nrow, 1, 1, nnz_hint).lanes;
    hipStream_t s = wa::as_stream(stream);
    switch (g) {
    case 1: return wa::launch_sweep<1>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    case 2: return wa::launch_sweep<2>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    case 4: return wa::launch_sweep<4>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    case 8: return wa::launch_sweep<8>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    case 16: return wa::launch_sweep<16>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    case 32: return wa::launch_sweep<32>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    default: return wa::launch_sweep<64>(nrow, offsets, columns, values, z, r, dinv, omega, z_out, s);
    }
}

// z += mult * (*scale) * out (network output back to the FP64 correction)
WA_API int warp_amd_unet_unpack_add(int64_t n, const float* out, const double* scale, double mult, double* z,
                                    void* stream)
{
    if (n <= 0)
        return wa::kOk;
    wa::unpack_add_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, out, scale, mult, z
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Loss and gradient of the one-step energy reduction (see energy_grad_kernel):
// sc (device) = {r.z, z.Az, r.e}; g = dL/dz; *loss (device) = L.
WA_API int warp_amd_energy_loss_grad(int64_t n, const double* r, const double* az, const double* sc, double* g,
                                     double* loss, void* stream)
{
    WA_REQUIRE(n > 0, "energy_loss_grad: empty vectors");
    wa::energy_grad_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, r, az, sc, g, loss
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// out32 = mult * (*scale) * x64
WA_API int warp_amd_scale_to_f32(int64_t n, const double* x, const double* scale, double mult, float* out,
                                 void* stream)
{
    if (n <= 0)
        return wa::kOk;
    wa::scale_to_f32_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, x, scale, mult, out
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// ---------------------------------------------------------------------------
// Fused flexible-CG steps (scalar FP32 matrix, FP64 vectors).  Workspace:
// warp_amd_fcg_fused_workspace_size(nrow) bytes of partial sums.
// ---------------------------------------------------------------------------

WA_API size_t warp_amd_fcg_fused_workspace_size(int nrow)
{
    return sizeof(double) * 2 * (size_t)(wa::ceil_div(nrow > 0 ? nrow : 1, 4) + 1);
}

#define WA_FCG_DISPATCH(FN, ...)                                                                               \
    do {                                                                                                       \
        WA_REQUIRE(nrow > 0, "fcg: empty matrix");                                                             \
        WA_REQUIRE(workspace && workspace_bytes >= warp_amd_fcg_fused_workspace_size(nrow),                    \
                   "fcg: workspace too small");                                                                \
        const int g = wa::default_bsr_shape(nrow, 1, 1, nnz_hint).lanes;                                       \
        double* partials = static_cast<double*>(workspace);                                                    \
        hipStream_t s = wa::as_stream(stream);                                                                 \
        switch (g) {                                                                                           \
        case 1: return wa::FcgLaunch<1>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);          \
        case 2: return wa::FcgLaunch<2>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);          \
        case 4: return wa::FcgLaunch<4>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);          \
        case 8: return wa::FcgLaunch<8>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);          \
        case 16: return wa::FcgLaunch<16>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);        \
        case 32: return wa::FcgLaunch<32>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);        \
        default: return wa::FcgLaunch<64>::FN(nrow, offsets, columns, values, __VA_ARGS__, partials, s);        \
        }                                                                                                      \
    } while (0)

// w = (float)(omega D^-1 r) (the first pre-smoothing sweep; fcg_update keeps it current)
WA_API int warp_amd_fcg_weight(int64_t n, const double* dinv, const double* r, double omega, float* w, void* stream)
{
    if (n <= 0)
        return wa::kOk;
    wa::fcg_weight_kernel<<<(unsigned)wa::ceil_div(n, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        n, dinv, r, omega, w
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// r1 = r - A w; *rr1 = r1.r1, *scale = rms(r1), *inv_scale = 1 / rms(r1)  (device)
WA_API int warp_amd_fcg_presmooth(int nrow, const int* offsets, const int* columns, const float* values,
                                  int64_t nnz_hint, const double* r, const float* w, double* r1, double* rr1,
                                  double* scale, double* inv_scale, void* workspace, size_t workspace_bytes,
                                  void* stream)
{
    WA_FCG_DISPATCH(pre, r, w, r1, rr1, scale, inv_scale);
}

// z2 = w + gain * (*scale) * out;  z = (float)(z2 + omega D^-1 (r - A z2));  *zq = z.q_prev  (device)
WA_API int warp_amd_fcg_postsmooth(int nrow, const int* offsets, const int* columns, const float* values,
                                   int64_t nnz_hint, const double* r, const double* dinv, double omega, const float* w,
                                   const float* out, const double* scale, double gain, const double* q_prev, float* z,
                                   double* zq, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_FCG_DISPATCH(post, r, dinv, omega, w, out, scale, gain, q_prev, z, zq);
}

// p = z - (zq / pq_prev) p_prev;  q = A p;  *pq = p.q, *pr = p.r  (device)
WA_API int warp_amd_fcg_direction_spmv(int nrow, const int* offsets, const int* columns, const float* values,
                                       int64_t nnz_hint, const float* z, const double* p_prev, const double* zq,
                                       const double* pq_prev, const double* r, double* p, double* q, double* pq,
                                       double* pr, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_FCG_DISPATCH(dir, z, p_prev, zq, pq_prev, r, p, q, pq, pr);
}
#undef WA_FCG_DISPATCH
