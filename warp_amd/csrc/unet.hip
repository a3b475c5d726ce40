// 3D U-Net building blocks for the learned flexible-CG preconditioner.
//
// No reference counterpart (the reference has no learned preconditioner);
// the Adam step follows warp/_src/optim/adam.py:30-47.  Fields on an n^3
// grid are channels-last FP32 arrays [voxel][channel], voxel = x + n (y + n z).
// Convolutions are 3x3x3, stride 1, zero padding, weights W[tap][ci][co] with
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1):  y[v, co] = b[co] + sum x[v + d, ci] W.
//
// FP32 kernels here serve training (forward, backward-data, backward-weight
// with a deterministic chunked reduction) and as the reference for the BF16
// MFMA inference kernel (unet_mfma.hip).
#include "common.h"

namespace wa {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void tap_offset(int t, int& dz, int& dy, int& dx)
{
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

// out[v, o] (+)= bias[o] + sum_t sum_r in[v +- d_t, r] * W(t, r, o)
// FLIP = false: forward, W(t, r, o) = W[(t*CR + r)*CO + o], neighbour v + d_t
// FLIP = true : backward data, W(t, r, o) = W[(t*CO + o)*CR + r], neighbour v - d_t
template <int CR, int CO, int COT, bool FLIP>
__global__ __launch_bounds__(kBlock) void conv3_kernel(
    int n, const float* __restrict__ in, const float* __restrict__ W, const float* __restrict__ bias,
    float* __restrict__ out, int relu, int accumulate
)
{
    const int64_t nvox = (int64_t)n * n * n;
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int o0 = blockIdx.y * COT;
    if (v >= nvox)
        return;
    const int x = (int)(v % n);
    const int y = (int)((v / n) % n);
    const int z = (int)(v / ((int64_t)n * n));
    float acc[COT];
#pragma unroll
    for (int j = 0; j < COT; ++j)
        acc[j] = (bias && o0 + j < CO) ? bias[o0 + j] : 0.0f;
    for (int t = 0; t < 27; ++t) {
        int dz, dy, dx;
        tap_offset(t, dz, dy, dx);
        if (FLIP) {
            dz = -dz;
            dy = -dy;
            dx = -dx;
        }
        const int nx = x + dx, ny = y + dy, nz = z + dz;
        if (nx < 0 || nx >= n || ny < 0 || ny >= n || nz < 0 || nz >= n)
            continue;
        const float* src = in + ((int64_t)nx + (int64_t)n * (ny + (int64_t)n * nz)) * CR;
#pragma unroll 4
        for (int r = 0; r < CR; ++r) {
            const float xv = src[r];
#pragma unroll
            for (int j = 0; j < COT; ++j) {
                const int o = o0 + j;
                if (o < CO) {
                    const float w = FLIP ? W[((int64_t)t * CO + o) * CR + r] : W[((int64_t)t * CR + r) * CO + o];
                    acc[j] = fmaf(xv, w, acc[j]);
                }
            }
        }
    }
    float* dst = out + v * CO + o0;
#pragma unroll
    for (int j = 0; j < COT; ++j) {
        if (o0 + j < CO) {
            float a = acc[j];
            if (accumulate)
                a += dst[j];
            dst[j] = relu ? fmaxf(a, 0.0f) : a;
        }
    }
}

// Per-chunk partial weight gradients: partial[((chunk*27 + t)*CI + ci)*CO + co]
// = sum_{v in chunk} x[v + d_t, ci] dy[v, co]; 16 x 16 (ci, co) tile per workgroup.
template <int CI, int CO>
__global__ __launch_bounds__(kBlock) void conv3_wgrad_kernel(
    int n, int chunk, const float* __restrict__ x, const float* __restrict__ dyv, float* __restrict__ partial
)
{
    __shared__ float xs[16][17];
    __shared__ float ds[16][17];
    const int64_t nvox = (int64_t)n * n * n;
    const int c = blockIdx.x;
    const int t = blockIdx.y;
    constexpr int kCoTiles = (CO + 15) / 16;
    const int ci0 = (blockIdx.z / kCoTiles) * 16;
    const int co0 = (blockIdx.z % kCoTiles) * 16;
    const int li = threadIdx.x / 16;  // ci within tile (compute); voxel (staging)
    const int lo = threadIdx.x % 16;  // co within tile (compute); channel (staging)
    int dz, dy, dx;
    tap_offset(t, dz, dy, dx);
    const int64_t v0 = (int64_t)c * chunk;
    const int64_t v1 = v0 + chunk < nvox ? v0 + chunk : nvox;
    float acc = 0.0f;
    for (int64_t vb = v0; vb < v1; vb += 16) {
        // stage 16 voxels x 16 channels of the shifted input and of dy
        const int64_t v = vb + li;
        float xv = 0.0f, dv = 0.0f;
        if (v < v1) {
            const int vx = (int)(v % n), vy = (int)((v / n) % n), vz = (int)(v / ((int64_t)n * n));
            const int nx = vx + dx, ny = vy + dy, nz = vz + dz;
            if (ci0 + lo < CI && nx >= 0 && nx < n && ny >= 0 && ny < n && nz >= 0 && nz < n)
                xv = x[((int64_t)nx + (int64_t)n * (ny + (int64_t)n * nz)) * CI + ci0 + lo];
            if (co0 + lo < CO)
                dv = dyv[v * CO + co0 + lo];
        }
        xs[li][lo] = xv;
        ds[li][lo] = dv;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k)
            acc = fmaf(xs[k][li], ds[k][lo], acc);
        __syncthreads();
    }
    const int ci = ci0 + li, co = co0 + lo;
    if (ci < CI && co < CO)
        partial[(((int64_t)c * 27 + t) * CI + ci) * CO + co] = acc;
}

// Per-chunk partial bias gradients: partial[chunk * CO + co] = sum_v dy[v, co].
__global__ __launch_bounds__(kBlock) void conv3_bgrad_kernel(
    int64_t nvox, int co_count, int chunk, const float* __restrict__ dyv, float* __restrict__ partial
)
{
    const int c = blockIdx.x;
    const int co = threadIdx.x;
    if (co >= co_count)
        return;
    const int64_t v0 = (int64_t)c * chunk;
    const int64_t v1 = v0 + chunk < nvox ? v0 + chunk : nvox;
    float acc = 0.0f;
    for (int64_t v = v0; v < v1; ++v)
        acc += dyv[v * co_count + co];
    partial[(int64_t)c * co_count + co] = acc;
}

// out[i] = sum over chunks of partial[chunk * count + i], in chunk order (FP64).
__global__ __launch_bounds__(kBlock) void sum_chunks_kernel(
    int64_t count, int nchunk, const float* __restrict__ partial, float* __restrict__ out
)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count)
        return;
    double a = 0.0;
    for (int c = 0; c < nchunk; ++c)
        a += (double)partial[(int64_t)c * count + i];
    out[i] = (float)a;
}

__global__ __launch_bounds__(kBlock) void relu_mask_kernel(
    int64_t count, const float* __restrict__ y, const float* __restrict__ dy, float* __restrict__ out
)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < count)
        out[i] = y[i] > 0.0f ? dy[i] : 0.0f;
}

// coarse voxel (cx, cy, cz) <-> fine children (2cx + a, 2cy + b, 2cz + c)
__global__ __launch_bounds__(kBlock) void pool2_kernel(int nf, int ch, const float* __restrict__ x, float* __restrict__ y)
{
    const int nc = nf / 2;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)nc * nc * nc * ch)
        return;
    const int c = (int)(i % ch);
    const int64_t p = i / ch;
    const int px = (int)(p % nc), py = (int)((p / nc) % nc), pz = (int)(p / ((int64_t)nc * nc));
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int fx = 2 * px + (k & 1), fy = 2 * py + ((k >> 1) & 1), fz = 2 * pz + (k >> 2);
        s += x[((int64_t)fx + (int64_t)nf * (fy + (int64_t)nf * fz)) * ch + c];
    }
    y[i] = s * 0.125f;
}

__global__ __launch_bounds__(kBlock) void pool2_bwd_kernel(
    int nf, int ch, const float* __restrict__ dy, float* __restrict__ dx, int accumulate
)
{
    const int nc = nf / 2;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)nf * nf * nf * ch)
        return;
    const int c = (int)(i % ch);
    const int64_t f = i / ch;
    const int fx = (int)(f % nf), fy = (int)((f / nf) % nf), fz = (int)(f / ((int64_t)nf * nf));
    const float g = dy[((int64_t)(fx / 2) + (int64_t)nc * (fy / 2 + (int64_t)nc * (fz / 2))) * ch + c] * 0.125f;
    dx[i] = accumulate ? dx[i] + g : g;
}

__global__ __launch_bounds__(kBlock) void up2_add_kernel(
    int nc, int ch, const float* __restrict__ x, const float* __restrict__ skip, float* __restrict__ y
)
{
    const int nf = 2 * nc;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)nf * nf * nf * ch)
        return;
    const int c = (int)(i % ch);
    const int64_t f = i / ch;
    const int fx = (int)(f % nf), fy = (int)((f / nf) % nf), fz = (int)(f / ((int64_t)nf * nf));
    const float v = x[((int64_t)(fx / 2) + (int64_t)nc * (fy / 2 + (int64_t)nc * (fz / 2))) * ch + c];
    y[i] = skip ? v + skip[i] : v;
}

__global__ __launch_bounds__(kBlock) void up2_bwd_kernel(int nc, int ch, const float* __restrict__ dy, float* __restrict__ dx)
{
    const int nf = 2 * nc;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)nc * nc * nc * ch)
        return;
    const int c = (int)(i % ch);
    const int64_t p = i / ch;
    const int px = (int)(p % nc), py = (int)((p / nc) % nc), pz = (int)(p / ((int64_t)nc * nc));
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int fx = 2 * px + (k & 1), fy = 2 * py + ((k >> 1) & 1), fz = 2 * pz + (k >> 2);
        s += dy[((int64_t)fx + (int64_t)nf * (fy + (int64_t)nf * fz)) * ch + c];
    }
    dx[i] = s;
}

// dout = (out - target) / count; partials of (out - target)^2
__global__ __launch_bounds__(kBlock) void mse_kernel(
    int64_t count, const float* __restrict__ out, const float* __restrict__ target, float* __restrict__ dout,
    double* __restrict__ partials
)
{
    __shared__ double lds[kBlock / kWave];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < count) {
        const float d = out[i] - target[i];
        dout[i] = d / (float)count;
        acc[0] = (double)d * (double)d;
    }
    write_block_partials<kBlock, 1>(acc, partials, lds);
}

__global__ __launch_bounds__(1024) void mse_finalize_kernel(const double* partials, int nparts, int64_t count, double* loss)
{
    __shared__ double lds[1024 / kWave];
    const double s = sum_partials<1024>(partials, nparts, lds);
    if (threadIdx.x == 0)
        *loss = 0.5 * s / (double)count;
}

// Adam (warp/_src/optim/adam.py:30-47): bias correction with t + 1, eps outside the sqrt.
__global__ __launch_bounds__(kBlock) void adam_kernel(
    int64_t count, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, float* __restrict__ p,
    float lr, float beta1, float beta2, float t, float eps
)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count)
        return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mhat = mi / (1.0f - powf(beta1, t + 1.0f));
    const float vhat = vi / (1.0f - powf(beta2, t + 1.0f));
    p[i] = p[i] - lr * mhat / (sqrtf(vhat) + eps);
}

// y = alpha * x + beta * y (gradient accumulation over a batch)
__global__ __launch_bounds__(kBlock) void axpby_f32_kernel(int64_t count, float alpha, const float* __restrict__ x,
                                                          float beta, float* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < count)
        y[i] = beta == 0.0f ? alpha * x[i] : fmaf(alpha, x[i], beta * y[i]);
}

// inp[v] = (r[v] * inv_scale, logk_std[v]); inv_scale read from device memory.
__global__ __launch_bounds__(kBlock) void pack_input_kernel(
    int64_t nvox, const double* __restrict__ r, const float* __restrict__ logk_std, const double* inv_scale,
    float* __restrict__ inp
)
{
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= nvox)
        return;
    inp[2 * v] = (float)(r[v] * (*inv_scale));
    inp[2 * v + 1] = logk_std[v];
}

// z[v] = scale * out[v] (FP64 result of the preconditioner)
__global__ __launch_bounds__(kBlock) void unpack_output_kernel(
    int64_t nvox, const float* __restrict__ out, const double* scale, double* __restrict__ z
)
{
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < nvox)
        z[v] = (double)out[v] * (*scale);
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------

unsigned grid1(int64_t count) { return (unsigned)ceil_div(count > 0 ? count : 1, kBlock); }

template <int CR, int CO, bool FLIP>
int launch_conv(int n, const float* in, const float* W, const float* b, float* out, int relu, int acc, hipStream_t s)
{
    constexpr int COT = CO < 16 ? CO : 16;
    const int64_t nvox = (int64_t)n * n * n;
    dim3 grid(grid1(nvox), (CO + COT - 1) / COT);
    conv3_kernel<CR, CO, COT, FLIP><<<grid, kBlock, 0, s>>>(n, in, W, b, out, relu, acc);
    WA_LAUNCH_CHECK();
    return kOk;
}

#define WA_CONV_PAIRS(X) X(2, 16) X(16, 32) X(32, 64) X(64, 64) X(64, 32) X(32, 16) X(16, 1) X(16, 16) X(1, 16)

int conv_fwd(int n, int cin, int cout, const float* x, const float* W, const float* b, float* y, int relu, int acc,
             hipStream_t s)
{
#define WA_FWD(CI, CO)                                                                                                \
    if (cin == CI && cout == CO)                                                                                      \
        return launch_conv<CI, CO, false>(n, x, W, b, y, relu, acc, s);
    WA_CONV_PAIRS(WA_FWD)
#undef WA_FWD
    set_error("conv3: unsupported channel pair %d -> %d", cin, cout);
    return kErrUnsupported;
}

int conv_bwd_data(int n, int cin, int cout, const float* dy, const float* W, float* dx, int acc, hipStream_t s)
{
#define WA_BWD(CI, CO)                                                                                                \
    if (cin == CI && cout == CO)                                                                                      \
        return launch_conv<CO, CI, true>(n, dy, W, nullptr, dx, 0, acc, s);
    WA_CONV_PAIRS(WA_BWD)
#undef WA_BWD
    set_error("conv3 backward: unsupported channel pair %d -> %d", cin, cout);
    return kErrUnsupported;
}

constexpr int kWgradChunk = 4096;

int wgrad_chunks(int n) { return (int)ceil_div((int64_t)n * n * n, kWgradChunk); }

template <int CI, int CO>
int launch_wgrad(int n, const float* x, const float* dy, float* partial, hipStream_t s)
{
    dim3 grid(wgrad_chunks(n), 27, ((CI + 15) / 16) * ((CO + 15) / 16));
    conv3_wgrad_kernel<CI, CO><<<grid, kBlock, 0, s>>>(n, kWgradChunk, x, dy, partial);
    WA_LAUNCH_CHECK();
    return kOk;
}

int conv_bwd_weight(int n, int cin, int cout, const float* x, const float* dy, float* partial, hipStream_t s)
{
#define WA_WG(CI, CO)                                                                                                 \
    if (cin == CI && cout == CO)                                                                                      \
        return launch_wgrad<CI, CO>(n, x, dy, partial, s);
    WA_CONV_PAIRS(WA_WG)
#undef WA_WG
    set_error("conv3 weight gradient: unsupported channel pair %d -> %d", cin, cout);
    return kErrUnsupported;
}

}  // namespace
}  // namespace wa

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// y = conv3(x) + b (ReLU when relu != 0); accumulate != 0 adds into y first.
WA_API int warp_amd_conv3_fwd_f32(int n, int cin, int cout, const float* x, const float* W, const float* b, float* y,
                                  int relu, int accumulate, void* stream)
{
    WA_REQUIRE(n > 0, "conv3: empty grid");
    return wa::conv_fwd(n, cin, cout, x, W, b, y, relu, accumulate, wa::as_stream(stream));
}

// dx (+)= conv3^T(dy): gradient of the convolution input.
WA_API int warp_amd_conv3_bwd_data_f32(int n, int cin, int cout, const float* dy, const float* W, float* dx,
                                       int accumulate, void* stream)
{
    WA_REQUIRE(n > 0, "conv3 backward: empty grid");
    return wa::conv_bwd_data(n, cin, cout, dy, W, dx, accumulate, wa::as_stream(stream));
}

WA_API size_t warp_amd_conv3_bwd_weight_workspace_size(int n, int cin, int cout)
{
    const size_t chunks = (size_t)wa::wgrad_chunks(n);
    return sizeof(float) * chunks * ((size_t)27 * cin * cout + cout) + 256;
}

// dW[t][ci][co] = sum_v x[v + d_t, ci] dy[v, co], db[co] = sum_v dy[v, co],
// each summed per 4096-voxel chunk and then over chunks in a fixed order.
WA_API int warp_amd_conv3_bwd_weight_f32(int n, int cin, int cout, const float* x, const float* dy, float* dW,
                                         float* db, void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(n > 0, "conv3 weight gradient: empty grid");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_conv3_bwd_weight_workspace_size(n, cin, cout),
               "conv3 weight gradient: workspace too small");
    WA_REQUIRE(cout <= wa::kBlock, "conv3 weight gradient: at most %d output channels", wa::kBlock);
    hipStream_t s = wa::as_stream(stream);
    const int nchunk = wa::wgrad_chunks(n);
    float* pw = static_cast<float*>(workspace);
    float* pb = pw + (size_t)nchunk * 27 * cin * cout;
    WA_TRY(wa::conv_bwd_weight(n, cin, cout, x, dy, pw, s));
    const int64_t wcount = (int64_t)27 * cin * cout;
    wa::sum_chunks_kernel<<<wa::grid1(wcount), wa::kBlock, 0, s>>>(wcount, nchunk, pw, dW);
    WA_LAUNCH_CHECK();
    const int64_t nvox = (int64_t)n * n * n;
    wa::conv3_bgrad_kernel<<<nchunk, wa::kBlock, 0, s>>>(nvox, cout, wa::kWgradChunk, dy, pb);
    WA_LAUNCH_CHECK();
    wa::sum_chunks_kernel<<<wa::grid1(cout), wa::kBlock, 0, s>>>(cout, nchunk, pb, db);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// out = dy * (y > 0)
WA_API int warp_amd_relu_mask_f32(int64_t count, const float* y, const float* dy, float* out, void* stream)
{
    wa::relu_mask_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(count, y, dy, out);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// y (n/2)^3 x ch = 2x2x2 average of x (n^3 x ch)
WA_API int warp_amd_pool2_f32(int n, int ch, const float* x, float* y, void* stream)
{
    WA_REQUIRE(n % 2 == 0, "pool2: grid side must be even");
    const int64_t count = (int64_t)(n / 2) * (n / 2) * (n / 2) * ch;
    wa::pool2_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(n, ch, x, y);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_pool2_bwd_f32(int n, int ch, const float* dy, float* dx, int accumulate, void* stream)
{
    WA_REQUIRE(n % 2 == 0, "pool2: grid side must be even");
    const int64_t count = (int64_t)n * n * n * ch;
    wa::pool2_bwd_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(n, ch, dy, dx, accumulate);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// y (2n)^3 x ch = nearest upsampling of x (n^3 x ch) + skip (nullable)
WA_API int warp_amd_up2_add_f32(int n, int ch, const float* x, const float* skip, float* y, void* stream)
{
    const int64_t count = (int64_t)8 * n * n * n * ch;
    wa::up2_add_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(n, ch, x, skip, y);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_up2_bwd_f32(int n, int ch, const float* dy, float* dx, void* stream)
{
    const int64_t count = (int64_t)n * n * n * ch;
    wa::up2_bwd_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(n, ch, dy, dx);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API size_t warp_amd_mse_workspace_size(int64_t count) { return sizeof(double) * (size_t)wa::grid1(count); }

// dout = (out - target) / count; *loss (device) = 0.5 mean((out - target)^2)
WA_API int warp_amd_mse_loss_f32(int64_t count, const float* out, const float* target, float* dout, double* loss,
                                 void* workspace, size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_mse_workspace_size(count), "mse: workspace too small");
    hipStream_t s = wa::as_stream(stream);
    const unsigned grid = wa::grid1(count);
    wa::mse_kernel<<<grid, wa::kBlock, 0, s>>>(count, out, target, dout, static_cast<double*>(workspace));
    WA_LAUNCH_CHECK();
    wa::mse_finalize_kernel<<<1, 1024, 0, s>>>(static_cast<double*>(workspace), (int)grid, count, loss);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// Adam step (replaces adam_step_kernel_float, warp/_src/optim/adam.py:30-47)
WA_API int warp_amd_adam_step_f32(int64_t count, const float* g, float* m, float* v, float* param, float lr,
                                  float beta1, float beta2, float t, float eps, void* stream)
{
    wa::adam_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(count, g, m, v, param, lr, beta1,
                                                                                  beta2, t, eps);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_axpby_f32(int64_t count, float alpha, const float* x, float beta, float* y, void* stream)
{
    if (count <= 0)
        return wa::kOk;
    wa::axpby_f32_kernel<<<wa::grid1(count), wa::kBlock, 0, wa::as_stream(stream)>>>(count, alpha, x, beta, y);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_unet_pack_input(int64_t nvox, const double* r, const float* logk_std, const double* inv_scale,
                                    float* inp, void* stream)
{
    wa::pack_input_kernel<<<wa::grid1(nvox), wa::kBlock, 0, wa::as_stream(stream)>>>(nvox, r, logk_std, inv_scale,
                                                                                      inp);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_unet_unpack_output(int64_t nvox, const float* out, const double* scale, double* z, void* stream)
{
    wa::unpack_output_kernel<<<wa::grid1(nvox), wa::kBlock, 0, wa::as_stream(stream)>>>(nvox, out, scale, z);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}
