// Mixed-precision pieces for FP64 iterative refinement around the FP32 CG,
// and the vector kernels of the flexible CG with the learned preconditioner.
//
// No reference counterpart: the reference solves in one precision
// (warp/_src/optim/linear.py).  Classical iterative refinement (e.g. Carson &
// Higham, SIAM J. Sci. Comput. 40(2), 2018): r = b - A x in FP64 with the FP32
// matrix promoted in-kernel, solve A d = r in FP32, x += d in FP64.
#include "bsr_core.h"

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
__global__ __launch_bounds__(kBlock) void fcg_update_kernel(
    int64_t n, const double* pr, const double* pq, const double* __restrict__ p, const double* __restrict__ q,
    double* __restrict__ x, double* __restrict__ r, double* __restrict__ partials
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
                               double* x, double* r, double* rr, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_fcg_update_workspace_size(n), "fcg_update: workspace too small");
    hipStream_t s = wa::as_stream(stream);
    const unsigned grid = (unsigned)wa::ceil_div(n, wa::kBlock);
    double* partials = static_cast<double*>(workspace);
    wa::fcg_update_kernel<<<grid, wa::kBlock, 0, s>>>(n, pr, pq, p, q, x, r, partials);
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
