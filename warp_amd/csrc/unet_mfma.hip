// BF16 MFMA implicit-GEMM 3x3x3 convolution for the learned preconditioner.
//
// No reference counterpart.  Same conventions as the FP32 kernels (unet.hip):
// channels-last fields [voxel][channel] on an m^3 grid, voxel = x + m (y + m z),
// weights W[tap][ci][co], tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding.
// Activations and weights are BF16 (round to nearest even), products are
// accumulated in FP32 by gfx950's v_mfma_f32_16x16x32_bf16:
//
//   C[co][voxel] = sum_k A[co][k] B[k][voxel],   k = tap * CI + ci
//
// (k zero-padded to a multiple of 32).  Operand lane layout (64 lanes,
// l = 16 q + n): A: lane holds row n, k = 8q..8q+7 of the 32-wide block;
// B: lane holds column n, the same 8 k; C: lane holds column n, rows
// 4q..4q+3.  With A = weights and B = input, a lane's B operand is 8
// consecutive channels of one voxel and tap (one 16-byte LDS read; for
// CI = 16 a block pairs two taps, lanes q < 2 reading the first), and its
// C fragment 4 consecutive output channels of one voxel (one 8-byte store).
// Weights are pre-swizzled into A-operand order (16 bytes per lane).
//
// Work split: a wavefront owns a 4x4x4 voxel cube; column n = (y%4) * 4 +
// (x%4) and the four MFMA column tiles g = 0..3 are the cube's z slices.  The
// 2x2x2 average pool of the output is then a sum over tile pairs (g, g+1)
// plus the lane pairs n^1 (x) and n^4 (y), done in registers.  Four cubes
// share one workgroup, one LDS copy of their input halo and one LDS copy of
// the weights of the workgroup's output channels (loaded once: workgroups
// are persistent and keep their channel group).
//
// Input channel counts 1 and 2 (first layer, backward of the last) fold
// several taps into one 8-wide k slot: k = tap * CI + ci.
#include "common.h"

#include <type_traits>

namespace wa {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mfma_bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;  // 4 wavefronts, one voxel cube each

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return (unsigned short)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two floats -> two BF16 (round to nearest even) in one v_cvt_pk_bf16_f32;
// lo in the low half.
__device__ __forceinline__ unsigned pack_bf16x2(float lo, float hi)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

// 32-wide k blocks of a CI-channel input (27 taps, zero-padded)
__host__ __device__ constexpr int k_blocks(int ci) { return (27 * ci + 31) / 32; }
__host__ __device__ constexpr int c_blocks(int co) { return (co + 15) / 16; }

// A-operand tiles: Wsw[(kb * CB + cb) * 64 + lane] = 8 BF16 of W(tap, c, co)
// for co = 16 cb + lane % 16, k = 32 kb + 8 (lane / 16) + j (j = 0..7),
// (tap, c) = (k / ci, k % ci); zero where tap >= 27 or co >= cout.
// transpose_flip: use W'(tap, ci, co) = W[26 - tap][co][ci] with W stored
// for a (co -> ci) layer, i.e. the backward-data convolution.
__global__ void swizzle_kernel(int ci, int co, int transpose_flip, const float* __restrict__ W, bf16x8* __restrict__ out,
                               int total)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total)
        return;
    const int CB = c_blocks(co);
    const int lane = idx % 64;
    const int cb = (idx / 64) % CB;
    const int kb = idx / (64 * CB);
    const int o = cb * 16 + lane % 16;
    bf16x8 v;
    for (int j = 0; j < 8; ++j) {
        const int k = kb * 32 + 8 * (lane / 16) + j;
        const int tap = k / ci, c = k % ci;
        float w = 0.0f;
        if (o < co && tap < 27)
            w = transpose_flip ? W[((size_t)(26 - tap) * co + o) * ci + c] : W[((size_t)tap * ci + c) * co + o];
        v[j] = (short)f2bf(w);
    }
    out[idx] = v;
}

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    return z;
}

__device__ __forceinline__ bf16x8 load8(const unsigned short* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ bf16x8 add8(bf16x8 a, bf16x8 b)
{
    unsigned r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        r[j] = pack_bf16x2(bf2f((unsigned short)a[2 * j]) + bf2f((unsigned short)b[2 * j]),
                           bf2f((unsigned short)a[2 * j + 1]) + bf2f((unsigned short)b[2 * j + 1]));
    return __builtin_bit_cast(bf16x8, r);
}

enum InMode { kInPlain = 0, kInUpAdd = 1 };
enum OutFlags { kOutRelu = 1, kOutPool = 2, kOutF32 = 4, kOutAccumulate = 8 };

// Persistent tile loops: workgroup b is dispatched to XCD b % kXcds (the
// MI355X command processor deals workgroups to its 8 XCDs round-robin), so
// each XCD walks its own contiguous chunk of tiles: tiles that share halo
// voxels then run on the same XCD and meet in its L2.  Any other mapping
// only costs locality, never correctness.  (b, nb): this workgroup's index
// and the number of workgroups walking the same tiles.
constexpr int kXcds = 8;

struct TileWalk {
    int t, end, step;
    __device__ __forceinline__ TileWalk(int ntiles, int b, int nb)
    {
        if (nb >= kXcds && nb % kXcds == 0) {
            const int xcd = b % kXcds;
            const int chunk = (ntiles + kXcds - 1) / kXcds;
            t = xcd * chunk + b / kXcds;
            end = min(ntiles, (xcd + 1) * chunk);
            step = nb / kXcds;
        } else {
            t = b;
            end = ntiles;
            step = nb;
        }
    }
    __device__ __forceinline__ explicit TileWalk(int ntiles) : TileWalk(ntiles, (int)blockIdx.x, (int)gridDim.x) {}
};

// Workgroup tile: 8 x 8 x 4 voxels = four 4x4x4 cubes (wave w: cube (w & 1,
// w >> 1) in x, y) and its 10 x 10 x 6 input halo.
constexpr int kTX = 8, kTY = 8, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;

// Bytes per halo voxel of the MFMA kernel: CI BF16 channels, plus 32 bytes of
// padding for CI = 64 (spreads the 16-byte operand reads of a wavefront's 16
// voxels over the banks); one 4-byte word for CI <= 2.  Measured at 128^3
// (us, padding 0 / 16 / 32): 16 -> 32 at 64^3 17.9 / 19.4 / 18.6, 32 -> 16 at
// 64^3 19.9 / 22.1 / 26.0, 64 -> 64 at 16^3 17.2 / 15.6 / 14.7, 64 -> 32 at
// 32^3 19.2 / 17.9 / 17.1.
__host__ __device__ constexpr int halo_bytes(int ci) { return ci <= 2 ? 4 : 2 * ci + (ci >= 64 ? 32 : 0); }

// Halo voxel offset of tap t relative to the window corner (dx, dy, dz = -1)
__host__ __device__ constexpr int tap_offset(int t) { return ((t / 9) * kHY + (t / 3) % 3) * kHX + t % 3; }

// LDS layout of conv3_mfma_kernel: [weights of CBG channel blocks | halo | 16 zero bytes | bias]
template <int CI, int CBG> struct ConvLds {
    static constexpr int kWeights = k_blocks(CI) * CBG * 1024;
    static constexpr int kHaloOff = kWeights;
    static constexpr int kZeroOff = kHaloOff + kHalo * halo_bytes(CI);
    static constexpr int kBiasOff = kZeroOff + 16;
    static constexpr int kBytes = kBiasOff + 64 * CBG;
};

// The 10 x 10 x 6 input halo of the tile at (X0, Y0, Z0), held in registers
// between its loads and its LDS stores, so that a workgroup's next tile is
// loaded while the MFMAs of the current one run.  Every load of a tile is
// issued before the first is used (at most 19 per thread); coordinates
// outside the grid load a clamped in-range voxel and store zeros (with
// kInUpAdd the input is x + upsample(coarse), rounded to BF16 once per voxel).
template <int CI, int IN> struct HaloStage {
    static constexpr int QPV = CI <= 2 ? 1 : CI / 8;  // 16-byte channel octets (or one word) per voxel
    static constexpr int TOTAL = kHalo * QPV;
    static constexpr int ITERS = (TOTAL + kBlock - 1) / kBlock;
    using T = typename std::conditional<(CI <= 2), unsigned, bf16x8>::type;
    T v[ITERS];
    T cv[IN == kInUpAdd ? ITERS : 1];
    bool in[ITERS];

    __device__ __forceinline__ void load(int m, int X0, int Y0, int Z0, const unsigned short* __restrict__ x,
                                         const unsigned short* __restrict__ coarse)
    {
        const int mh = m / 2;
#pragma unroll
        for (int u = 0; u < ITERS; ++u) {
            const int c = min(u * kBlock + (int)threadIdx.x, TOTAL - 1);
            const int hv = c / QPV, od = c % QPV;
            const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
            const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
            in[u] = gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m;
            const int cx = min(max(gx, 0), m - 1), cy = min(max(gy, 0), m - 1), cz = min(max(gz, 0), m - 1);
            const size_t g = (size_t)(cz * m + cy) * m + cx;
            if constexpr (CI <= 2) {
                v[u] = CI == 2 ? reinterpret_cast<const unsigned*>(x)[g] : (unsigned)x[g];
            } else {
                v[u] = load8(x + g * CI + 8 * od);
                if constexpr (IN == kInUpAdd)
                    cv[u] = load8(coarse + ((size_t)((cz >> 1) * mh + (cy >> 1)) * mh + (cx >> 1)) * CI + 8 * od);
            }
        }
    }

    __device__ __forceinline__ void store(unsigned char* halo) const
    {
#pragma unroll
        for (int u = 0; u < ITERS; ++u) {
            const int c = u * kBlock + threadIdx.x;
            if (c < TOTAL) {
                if constexpr (CI <= 2) {
                    reinterpret_cast<unsigned*>(halo)[c] = in[u] ? v[u] : 0u;
                } else {
                    bf16x8 w = v[u];
                    if constexpr (IN == kInUpAdd)
                        w = add8(w, cv[u]);
                    *reinterpret_cast<bf16x8*>(halo + (c / QPV) * halo_bytes(CI) + 16 * (c % QPV)) =
                        in[u] ? w : zero8();
                }
            }
        }
    }
};

// x: input [m^3][CI] BF16; with kInUpAdd the input is x[v] + coarse[v / 2]
// (coarse: [(m/2)^3][CI]), rounded to BF16.  Output: y BF16 [m^3][CO] or, with
// kOutF32, yf FP32 [m^3][CO] (kOutAccumulate adds to yf); pooled BF16
// [(m/2)^3][CO] with kOutPool.
// A workgroup computes CBG of the CO / 16 output-channel blocks: channel
// group blockIdx.x % (CB / CBG), whose weights it holds in LDS for all of its
// tiles; the groups split the chip so that coarse grids still launch enough
// workgroups.  The k loop is fully unrolled: every LDS address is a
// compile-time offset from one per-lane base, so the compiler can issue the
// operand reads of later blocks while the MFMAs of earlier ones run.
// Minimum waves per SIMD the compiler must fit in registers: the 2 -> 16
// input layer at 6 (96 -> 69 registers; 30.8 -> 28.6 us at 128^3), the others
// as the compiler chooses (a bound of 3 on the 16 -> 32 layer measured the same).
template <int CI, int CO> constexpr int conv_min_waves() { return (CI == 2 && CO == 16) ? 6 : 1; }

template <int CI, int CO, int IN, int CBG>
__global__ __launch_bounds__(kBlock, (conv_min_waves<CI, CO>())) void conv3_mfma_kernel(
    int m, const unsigned short* __restrict__ x, const unsigned short* __restrict__ coarse,
    const bf16x8* __restrict__ wsw, const float* __restrict__ bias, int flags, unsigned short* __restrict__ y,
    float* __restrict__ yf, unsigned short* __restrict__ pooled
)
{
    constexpr int KB = k_blocks(CI);
    constexpr int CB = c_blocks(CO);
    constexpr int HB = halo_bytes(CI);
    constexpr int SLICE = kHX * kHY;  // halo voxels per z slice
    using L = ConvLds<CI, CBG>;
    static_assert(CB % CBG == 0, "channel blocks per workgroup must divide the channel blocks");
    static_assert(L::kBytes <= 160 * 1024, "LDS budget");
    __shared__ __attribute__((aligned(16))) unsigned char smem[L::kBytes];
    unsigned char* halo = smem + L::kHaloOff;

    constexpr int nsplit = CB / CBG;
    const int group = (int)blockIdx.x % nsplit;
    const int cb0 = group * CBG;
    // the workgroup's weights, once: [kb][cbl][lane] 16 bytes each, every
    // load issued before the first store
    {
        constexpr int kW = KB * CBG * 64;
        constexpr int kWIters = (kW + kBlock - 1) / kBlock;
        bf16x8 wv[kWIters];
#pragma unroll
        for (int u = 0; u < kWIters; ++u) {
            const int i = min(u * kBlock + (int)threadIdx.x, kW - 1);
            const int ln = i % 64, cbl = (i / 64) % CBG, kb = i / (64 * CBG);
            wv[u] = wsw[(kb * CB + cb0 + cbl) * 64 + ln];
        }
#pragma unroll
        for (int u = 0; u < kWIters; ++u) {
            const int i = u * kBlock + threadIdx.x;
            if (i < kW)
                reinterpret_cast<bf16x8*>(smem)[i] = wv[u];
        }
    }
    if (threadIdx.x == 0)
        *reinterpret_cast<bf16x8*>(smem + L::kZeroOff) = zero8();

    const int tiles_x = (m + kTX - 1) / kTX;
    const int tiles_xy = tiles_x * tiles_x;
    const int ntiles = tiles_xy * (m / kTZ);
    const int mh = m / 2;
    const int lane = threadIdx.x % 64;
    const int w = threadIdx.x / 64;
    const int n = lane % 16, q = lane / 16;
    const int lx = 4 * (w & 1) + (n & 3);  // position in the tile
    const int ly = 4 * (w >> 1) + (n >> 2);
    const int lane_base = ly * kHX + lx;  // halo voxel of the window corner in slice 0
    const bf16x8* wl = reinterpret_cast<const bf16x8*>(smem) + lane;
    // the workgroup's bias values, in LDS: a global load used in the tile
    // loop makes the epilogue wait for every load in flight (the counter is in
    // order), the next tile's halo included
    float* bias_lds = reinterpret_cast<float*>(smem + L::kBiasOff);
    if (threadIdx.x < 16 * CBG) {
        const int o = cb0 * 16 + (int)threadIdx.x;
        bias_lds[threadIdx.x] = (bias && o < CO) ? bias[o] : 0.0f;
    }

    auto origin = [&](int t, int& X0, int& Y0, int& Z0) {
        X0 = (t % tiles_x) * kTX;
        Y0 = ((t / tiles_x) % tiles_x) * kTY;
        Z0 = (t / tiles_xy) * kTZ;
    };
    TileWalk walk(ntiles, (int)blockIdx.x / nsplit, (int)gridDim.x / nsplit);
    HaloStage<CI, IN> stage;
    if (walk.t < walk.end) {
        int X0, Y0, Z0;
        origin(walk.t, X0, Y0, Z0);
        stage.load(m, X0, Y0, Z0, x, coarse);
    }
    for (; walk.t < walk.end; walk.t += walk.step) {
    int X0, Y0, Z0;
    origin(walk.t, X0, Y0, Z0);
    const int vx = X0 + lx, vy = Y0 + ly, vz0 = Z0;
    const bool valid = (X0 + 4 * (w & 1)) < m && (Y0 + 4 * (w >> 1)) < m;  // whole cube in range

    __syncthreads();  // weights stored / the previous tile's LDS reads are done
    stage.store(halo);
    __syncthreads();
    if (walk.t + walk.step < walk.end) {  // the next tile's halo, in flight during the MFMAs
        int X1, Y1, Z1;
        origin(walk.t + walk.step, X1, Y1, Z1);
        stage.load(m, X1, Y1, Z1, x, coarse);
    }

    f32x4 acc[CBG][4];
#pragma unroll
    for (int cb = 0; cb < CBG; ++cb)
#pragma unroll
        for (int g = 0; g < 4; ++g)
            acc[cb][g] = f32x4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        mfma_bf16x8 a[CBG];
#pragma unroll
        for (int cb = 0; cb < CBG; ++cb)
            a[cb] = __builtin_bit_cast(mfma_bf16x8, wl[(kb * CBG + cb) * 64]);
        mfma_bf16x8 b[4];
        if constexpr (CI >= 16) {
            // k = 32 kb + 8 q .. +7: one tap, 8 consecutive channels (lane-dependent tap only for CI = 16)
            const int k0 = 32 * kb + 8 * q;
            const int tap = k0 / CI, c0 = k0 % CI;
            const bool pad = tap >= 27;
            const unsigned char* src = pad ? smem + L::kZeroOff : halo + (lane_base + tap_offset(tap)) * HB + 2 * c0;
            const int gstep = pad ? 0 : SLICE * HB;
#pragma unroll
            for (int g = 0; g < 4; ++g)
                b[g] = __builtin_bit_cast(mfma_bf16x8, *reinterpret_cast<const bf16x8*>(src + g * gstep));
        } else {
            // k = 32 kb + 8 q + j: tap (k / CI), channel k % CI; one 4-byte halo word per voxel
            constexpr int TPS = 8 / CI;  // taps per 8-wide slot
            const unsigned* hw = reinterpret_cast<const unsigned*>(halo);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                bf16x8 v;
#pragma unroll
                for (int i = 0; i < TPS; ++i) {
                    const int tap = (32 * kb + 8 * q) / CI + i;
                    const unsigned u = tap < 27 ? hw[lane_base + tap_offset(tap) + g * SLICE] : 0u;
                    if constexpr (CI == 2) {
                        v[2 * i] = (short)(u & 0xffffu);
                        v[2 * i + 1] = (short)(u >> 16);
                    } else {
                        v[i] = (short)(u & 0xffffu);
                    }
                }
                b[g] = __builtin_bit_cast(mfma_bf16x8, v);
            }
        }
#pragma unroll
        for (int cb = 0; cb < CBG; ++cb)
#pragma unroll
            for (int g = 0; g < 4; ++g)
                acc[cb][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb], b[g], acc[cb][g], 0, 0, 0);
    }
    if (!valid)
        continue;  // whole wavefront (only when m < 8); still joins the next tile's barriers

    // epilogue: lane holds output channels 16 cb + 4 q + i of voxel (vx, vy, vz0 + g)
    const bool relu = flags & kOutRelu;
#pragma unroll
    for (int cb = 0; cb < CBG; ++cb) {
        const int o0 = (cb0 + cb) * 16 + 4 * q;
        const float4 bsv = *reinterpret_cast<const float4*>(bias_lds + cb * 16 + 4 * q);
        float psum[2][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const size_t v = (size_t)((vz0 + g) * m + vy) * m + vx;
            float val[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                val[i] = acc[cb][g][i] + (i == 0 ? bsv.x : i == 1 ? bsv.y : i == 2 ? bsv.z : bsv.w);
                if (relu)
                    val[i] = fmaxf(val[i], 0.0f);
            }
            if (flags & kOutF32) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (o0 + i < CO) {
                        float* d = yf + v * CO + o0 + i;
                        *d = (flags & kOutAccumulate) ? *d + val[i] : val[i];
                    }
            } else {
                const unsigned st[2] = {pack_bf16x2(val[0], val[1]), pack_bf16x2(val[2], val[3])};
#pragma unroll
                for (int i = 0; i < 4; ++i)  // pool what was stored
                    val[i] = bf2f((unsigned short)(st[i / 2] >> (16 * (i % 2))));
                if constexpr (CO % 16 == 0)
                    *reinterpret_cast<bf16x4*>(y + v * CO + o0) = __builtin_bit_cast(bf16x4, st);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                psum[g / 2][i] = (g % 2) ? psum[g / 2][i] + val[i] : val[i];
        }
        if (flags & kOutPool) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float sv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    sv[i] = psum[h][i];
                    sv[i] += __shfl_xor(sv[i], 1);
                    sv[i] += __shfl_xor(sv[i], 4);
                }
                const unsigned st[2] = {pack_bf16x2(sv[0] * 0.125f, sv[1] * 0.125f),
                                        pack_bf16x2(sv[2] * 0.125f, sv[3] * 0.125f)};
                if ((n & 1) == 0 && (n & 4) == 0) {
                    const size_t pv = (size_t)(((vz0 >> 1) + h) * mh + (vy >> 1)) * mh + (vx >> 1);
                    if constexpr (CO % 16 == 0)
                        *reinterpret_cast<bf16x4*>(pooled + pv * CO + o0) = __builtin_bit_cast(bf16x4, st);
                }
            }
        }
    }
    }  // tiles
}

// Persistent launch: as many workgroups as can be resident at once (from the
// occupancy of this kernel, queried once), each looping over tiles.
template <auto Kernel, typename... Args>
void launch_tiles(int ntiles, hipStream_t s, Args... args)
{
    static const int per_cu = [] {  // one cache per kernel (Kernel is a template argument)
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, Kernel, kBlock, 0) != hipSuccess || nb < 1)
            nb = 1;
        return nb;
    }();
    const int slots = per_cu * device_cu_count();
    const unsigned grid = (unsigned)(ntiles < slots ? ntiles : slots);
    Kernel<<<grid, kBlock, 0, s>>>(args...);
}

// One output channel (the network's last layer), streamed through z: a
// workgroup owns 16 x 16 columns and kCo1Z consecutive output slices, and
// walks the input planes z' = Z0 - 1 .. Z0 + kCo1Z one at a time.  Per plane
// one MFMA GEMM forms, for every column (x, y),
//
//   P_dz(x, y) = sum_{dy, dx, c} W[dz][dy][dx][c] in[x + dx, y + dy, z'][c],   dz = 0, 1, 2
//
// (rows = dz, padded to 16 with zero weights; columns = 16 x positions of one
// y row; k = (dy, dx, c), 144 deep, padded to 5 blocks of 32), and the
// plane's three results go to the outputs z' + 1, z' and z' - 1 in three
// rolling FP32 accumulators per column: out = bias + P_0(z - 1) + P_1(z) +
// P_2(z + 1).  Only one 18 x 18 plane (its halo included) is in LDS at a
// time, double-buffered; the loads of the plane two steps ahead are in
// flight in registers.  (Earlier forms, at 128^3: v_dot2_f32_bf16 per voxel
// and tap, 39.3 us, bound by its 216 dot products per output; MFMA products
// per voxel and tap exchanged through LDS, 45.7 us, bound by its three
// barriers per plane.)
constexpr int kCo1T = 16;             // columns per side of a workgroup
constexpr int kCo1H = kCo1T + 2;      // halo plane side
// Output slices per workgroup: at 128^3, 12 gives 8 x 8 x 11 = 704 tiles for
// the 768 slots of three workgroups per CU (30.7 us; 8 slices: 1024 tiles in
// two uneven waves, 32.3 us; 16 slices: 512 tiles, 32.4 us).
constexpr int kCo1Z = 12;
constexpr int kCo1VB = 48;            // LDS bytes per voxel: 16 BF16 channels + 16 bytes of padding
constexpr int kCo1KB = 5;             // k blocks of 32: (dy, dx, c) = 144, padded to 160

// One halo plane in registers between its loads and its LDS stores.
template <int IN> struct Co1Plane {
    static constexpr int CI = 16;
    static constexpr int kOct = kCo1H * kCo1H * 2;            // 16-byte octets per plane
    static constexpr int kIters = (kOct + kBlock - 1) / kBlock;  // 3 per thread
    bf16x8 v[kIters], cv[IN == kInUpAdd ? kIters : 1];
    bool in[kIters];

    __device__ __forceinline__ void load(int m, int X0, int Y0, int gz, const unsigned short* __restrict__ x,
                                         const unsigned short* __restrict__ coarse)
    {
        const int mh = m / 2;
        const int cz = min(max(gz, 0), m - 1);
#pragma unroll
        for (int u = 0; u < kIters; ++u) {
            const int c = min(u * kBlock + (int)threadIdx.x, kOct - 1);
            const int hv = c / 2, od = c % 2;
            const int gx = X0 - 1 + hv % kCo1H, gy = Y0 - 1 + hv / kCo1H;
            in[u] = gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m;
            const int cx = min(max(gx, 0), m - 1), cy = min(max(gy, 0), m - 1);
            v[u] = load8(x + ((size_t)(cz * m + cy) * m + cx) * CI + 8 * od);
            if constexpr (IN == kInUpAdd)
                cv[u] = load8(coarse + ((size_t)((cz >> 1) * mh + (cy >> 1)) * mh + (cx >> 1)) * CI + 8 * od);
        }
    }

    __device__ __forceinline__ void store(unsigned char* dst) const
    {
#pragma unroll
        for (int u = 0; u < kIters; ++u) {
            const int c = u * kBlock + threadIdx.x;
            if (c < kOct) {
                bf16x8 w = v[u];
                if constexpr (IN == kInUpAdd)
                    w = add8(w, cv[u]);
                *reinterpret_cast<bf16x8*>(dst + (c / 2) * kCo1VB + 16 * (c % 2)) = in[u] ? w : zero8();
            }
        }
    }
};

// (at least 3 waves per SIMD: 172 -> 166 registers, 37.0 -> 32.0 us at 128^3)
template <int IN>
__global__ __launch_bounds__(kBlock, 3) void conv3_co1_stream_kernel(int m, const unsigned short* __restrict__ x,
                                                                 const unsigned short* __restrict__ coarse,
                                                                 const unsigned* __restrict__ wpk,
                                                                 const float* __restrict__ bias, int flags,
                                                                 float* __restrict__ yf)
{
    constexpr int CI = 16;
    constexpr int kPlane = kCo1H * kCo1H;            // 324 voxels
    constexpr int kRowsPerWave = kCo1T / (kBlock / kWave);  // 4 y rows of 16 columns per wave
    __shared__ __attribute__((aligned(16))) unsigned char plane[2][kPlane * kCo1VB];
    const int tiles_x = (m + kCo1T - 1) / kCo1T;
    const int tiles_z = (m + kCo1Z - 1) / kCo1Z;
    const int ntiles = tiles_x * tiles_x * tiles_z;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int ln = lane % 16, lq = lane / 16;  // ln: column (x), lq: k slot
    const float b0 = bias ? bias[0] : 0.0f;
    // A operands (constant): row ln = dz (< 3), k = 32 kb + 8 lq .. + 7 -> (dy, dx) = k / 16, channels k % 16
    mfma_bf16x8 wa[kCo1KB];
    // B operand offsets within a plane (bytes, relative to the column's corner voxel); -1: padding slot
    int boff[kCo1KB];
#pragma unroll
    for (int kb = 0; kb < kCo1KB; ++kb) {
        const int k0 = 32 * kb + 8 * lq;
        const int t9 = k0 / 16, c0 = k0 % 16;
        unsigned wv[4] = {0u, 0u, 0u, 0u};
        if (ln < 3 && t9 < 9) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                wv[j] = wpk[(ln * 9 + t9) * (CI / 2) + c0 / 2 + j];
        }
        wa[kb] = __builtin_bit_cast(mfma_bf16x8, wv);
        boff[kb] = t9 < 9 ? ((t9 / 3) * kCo1H + t9 % 3) * kCo1VB + 2 * c0 : -1;
    }

    for (TileWalk walk(ntiles); walk.t < walk.end; walk.t += walk.step) {
    const int X0 = (walk.t % tiles_x) * kCo1T;
    const int Y0 = ((walk.t / tiles_x) % tiles_x) * kCo1T;
    const int Z0 = (walk.t / (tiles_x * tiles_x)) * kCo1Z;
    const int nz = min(kCo1Z, m - Z0);
    const int np = nz + 2;  // planes z' = Z0 - 1 .. Z0 + nz (at least 3)

    // acc[j][0]: output z' - 1 of row j (its last plane), [1]: z', [2]: z' + 1 (its first plane);
    // meaningful in the lanes lq == 0 (C rows 0..2 = dz)
    float acc[kRowsPerWave][3];
#pragma unroll
    for (int j = 0; j < kRowsPerWave; ++j)
        acc[j][0] = acc[j][1] = acc[j][2] = b0;
    auto step = [&](Co1Plane<IN>& regs, int p) {
        unsigned char* cur = plane[p & 1];
        regs.store(cur);
        __syncthreads();  // the plane is complete; everyone is done with the slot stored next
        if (p + 2 < np)
            regs.load(m, X0, Y0, Z0 + 1 + p, x, coarse);  // plane p + 2
        const int zo = Z0 - 2 + p;  // output completed by this plane
#pragma unroll
        for (int j = 0; j < kRowsPerWave; ++j) {
            const int ty = wave * kRowsPerWave + j;
            const unsigned char* col = cur + (ty * kCo1H + ln) * kCo1VB;
            f32x4 c4 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < kCo1KB; ++kb) {
                const bf16x8 bv = boff[kb] >= 0 ? *reinterpret_cast<const bf16x8*>(col + boff[kb]) : zero8();
                c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[kb], __builtin_bit_cast(mfma_bf16x8, bv), c4, 0, 0, 0);
            }
            // plane z' is tap dz of output z' + 1 - dz
            acc[j][2] += c4[0];
            acc[j][1] += c4[1];
            acc[j][0] += c4[2];
            const int vx = X0 + ln, vy = Y0 + ty;
            if (lq == 0 && p >= 2 && vx < m && vy < m) {
                float r = acc[j][0];
                if (flags & kOutRelu)
                    r = fmaxf(r, 0.0f);
                float* dst = yf + (size_t)(zo * m + vy) * m + vx;
                *dst = (flags & kOutAccumulate) ? *dst + r : r;
            }
            acc[j][0] = acc[j][1];
            acc[j][1] = acc[j][2];
            acc[j][2] = b0;
        }
    };
    Co1Plane<IN> ra, rb;
    ra.load(m, X0, Y0, Z0 - 1, x, coarse);
    rb.load(m, X0, Y0, Z0, x, coarse);
    for (int p = 0; p < np; p += 2) {
        step(ra, p);
        if (p + 1 < np)
            step(rb, p + 1);
    }
    __syncthreads();  // the last planes' reads are done before the next tile stores
    }  // tiles
}

// wpk[tap][c / 2] = (W(tap, c, 0), W(tap, c + 1, 0)) as a BF16 pair (low, high)
__global__ void pack_co1_kernel(int ci, int transpose_flip, const float* __restrict__ W, unsigned* __restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 27 * (ci / 2))
        return;
    const int tap = idx / (ci / 2), c = 2 * (idx % (ci / 2));
    // (ci -> 1): W[tap][c][0];  transposed (layer 1 -> ci): W[26 - tap][0][c]
    const float w0 = transpose_flip ? W[(size_t)(26 - tap) * ci + c] : W[(size_t)tap * ci + c];
    const float w1 = transpose_flip ? W[(size_t)(26 - tap) * ci + c + 1] : W[(size_t)tap * ci + c + 1];
    out[idx] = (unsigned)f2bf(w0) | ((unsigned)f2bf(w1) << 16);
}

template <int CI>
int launch_co1(int m, int in_mode, const unsigned short* x, const unsigned short* coarse, const void* wpk,
               const float* bias, int flags, float* yf, hipStream_t s)
{
    static_assert(CI == 16, "the streamed one-output kernel takes 16 input channels");
    const int txy = (m + kCo1T - 1) / kCo1T;
    const unsigned* w = static_cast<const unsigned*>(wpk);
    const int ntiles = txy * txy * ((m + kCo1Z - 1) / kCo1Z);
    if (in_mode == kInUpAdd)
        launch_tiles<conv3_co1_stream_kernel<kInUpAdd>>(ntiles, s, m, x, coarse, w, bias, flags, yf);
    else
        launch_tiles<conv3_co1_stream_kernel<kInPlain>>(ntiles, s, m, x, coarse, w, bias, flags, yf);
    WA_LAUNCH_CHECK();
    return kOk;
}

template <int CI, int CO, int IN, int CBG>
void launch_conv_groups(int m, int tiles, const unsigned short* x, const unsigned short* coarse, const bf16x8* w,
                        const float* bias, int flags, unsigned short* y, float* yf, unsigned short* pooled,
                        hipStream_t s)
{
    constexpr auto kernel = conv3_mfma_kernel<CI, CO, IN, CBG>;
    constexpr int nsplit = c_blocks(CO) / CBG;
    static const int per_cu = [] {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, 0) != hipSuccess || nb < 1)
            nb = 1;
        return nb;
    }();
    // every channel group gets the same number of workgroups (each tile is
    // computed exactly once per group, also with kOutAccumulate)
    const int slots = per_cu * device_cu_count();
    int grid = tiles * nsplit < slots ? tiles * nsplit : slots;
    grid = grid / nsplit * nsplit;
    if (grid < nsplit)
        grid = nsplit;
    kernel<<<(unsigned)grid, kBlock, 0, s>>>(m, x, coarse, w, bias, flags, y, yf, pooled);
}

template <int CI, int CO, int IN>
void launch_conv_mode(int m, const unsigned short* x, const unsigned short* coarse, const bf16x8* w,
                      const float* bias, int flags, unsigned short* y, float* yf, unsigned short* pooled, hipStream_t s)
{
    constexpr int CB = c_blocks(CO);
    const int txy = (m + kTX - 1) / kTX;
    const int tiles = txy * txy * (m / kTZ);
    // all channel blocks in one workgroup when the spatial tiles alone fill
    // the chip and their weights fit in LDS beside the halo; else one block each
    constexpr bool kCanGroup = CB > 1 && ConvLds<CI, CB>::kBytes <= 160 * 1024;
    if constexpr (kCanGroup) {
        if (tiles >= 1024) {
            launch_conv_groups<CI, CO, IN, CB>(m, tiles, x, coarse, w, bias, flags, y, yf, pooled, s);
            return;
        }
    }
    launch_conv_groups<CI, CO, IN, 1>(m, tiles, x, coarse, w, bias, flags, y, yf, pooled, s);
}

template <int CI, int CO>
int launch_conv(int m, int in_mode, const unsigned short* x, const unsigned short* coarse, const void* wsw,
                const float* bias, int flags, unsigned short* y, float* yf, unsigned short* pooled, hipStream_t s)
{
    const bf16x8* w = static_cast<const bf16x8*>(wsw);
    if (in_mode == kInUpAdd) {
        if constexpr (CI > 2)
            launch_conv_mode<CI, CO, kInUpAdd>(m, x, coarse, w, bias, flags, y, yf, pooled, s);
    } else {
        launch_conv_mode<CI, CO, kInPlain>(m, x, coarse, w, bias, flags, y, yf, pooled, s);
    }
    WA_LAUNCH_CHECK();
    return kOk;
}

__global__ void f32_to_bf16_kernel(int64_t count, const float* __restrict__ in, unsigned short* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count)
        out[i] = f2bf(in[i]);
}

__global__ void bf16_to_f32_kernel(int64_t count, const unsigned short* __restrict__ in, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count)
        out[i] = bf2f(in[i]);
}

// (r * inv_scale, logk) -> BF16 [nvox][2]
__global__ void pack_input_bf16_kernel(int64_t nvox, const double* __restrict__ r, const float* __restrict__ logk,
                                       const double* inv_scale, unsigned short* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nvox)
        return;
    const double s = *inv_scale;
    const unsigned lo = f2bf((float)(r[i] * s));
    const unsigned hi = f2bf(logk[i]);
    reinterpret_cast<unsigned*>(out)[i] = lo | (hi << 16);
}

// The same, four voxels per thread in 16-byte loads and stores (r, logk and
// out 16-byte aligned, nvox a multiple of 4).
__global__ void pack_input_bf16_x4_kernel(int64_t nquad, const double* __restrict__ r,
                                          const float* __restrict__ logk, const double* inv_scale,
                                          unsigned short* __restrict__ out)
{
    typedef double D2 __attribute__((ext_vector_type(2)));
    typedef float F4 __attribute__((ext_vector_type(4)));
    typedef unsigned U4 __attribute__((ext_vector_type(4)));
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nquad)
        return;
    const double s = *inv_scale;
    const D2 r01 = reinterpret_cast<const D2*>(r)[2 * q];
    const D2 r23 = reinterpret_cast<const D2*>(r)[2 * q + 1];
    const F4 k = reinterpret_cast<const F4*>(logk)[q];
    const double rv[4] = {r01.x, r01.y, r23.x, r23.y};
    U4 o;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        o[e] = (unsigned)f2bf((float)(rv[e] * s)) | ((unsigned)f2bf(k[e]) << 16);
    reinterpret_cast<U4*>(out)[q] = o;
}

// ---------------------------------------------------------------------------
// Weight and bias gradients on MFMA (training):
//
//   dW[t][ci][co] = sum_v x~[v + d_t][ci] dy[v][co],   db[co] = sum_v dy[v][co]
//
// with x~ the layer input as the forward saw it (x, or x + upsample(coarse)
// rounded to BF16) and dy the FP32 output gradient, rounded to BF16 for the
// MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  GEMM per workgroup
// tile of 8 x 8 x 4 voxels: C[g][co] += A[g][k] B[k][co], k = 32 voxels (four
// x rows of 8), rows g = (tap, input channel) with g = t * CI + ci, and row
// 27 * CI a row of ones whose C row is db.  A lane's A operand is the 8
// voxels of one x row of the input shifted by its row's tap; three copies of
// the halo pre-shifted by dx = -1, 0, 1 (copy s holds halo x = s .. s + 7)
// keep that read 16-byte aligned.  Its B operand is the same 8 voxels of dy
// for one output channel (dy staged channel-major).  Rows are cut in blocks
// of 16: for CI >= 16 one block per tap and 16-channel group (workgroups
// split the channel groups) plus the ones block in group 0; for CI <= 2 the
// 27 * CI + 1 rows in order.  The four waves of a workgroup take every fourth
// row block.  Workgroups are persistent over their tiles, sum in registers,
// and write one partial [27 * CI + 1][CO_pad] each; wgrad_sum_kernel adds
// the partials in a fixed order.
constexpr int kWgXS = 6 * 10 * 8 * 2 + 16;  // bytes per (copy, channel) of the shifted halo (+16 padding)
constexpr int kWgDS = 32 * 8 * 2 + 16;       // bytes per output channel of the dy tile (+16 padding)
constexpr int kWgMaxSlots = 256;             // workgroups per channel group

__host__ __device__ constexpr int wg_channels(int ci) { return ci >= 16 ? 16 : ci; }
__host__ __device__ constexpr int wg_groups(int ci) { return ci >= 16 ? ci / 16 : 1; }
__host__ __device__ constexpr int wg_row_blocks(int ci) { return ci >= 16 ? 28 : (27 * ci + 1 + 15) / 16; }
__host__ __device__ constexpr int wg_rows(int ci) { return 27 * ci + 1; }

template <int CI, int CO, int IN>
__global__ __launch_bounds__(kBlock) void conv3_wgrad_mfma_kernel(int m, const unsigned short* __restrict__ x,
                                                                  const unsigned short* __restrict__ coarse,
                                                                  const float* __restrict__ dy, int nslots,
                                                                  float* __restrict__ partial)
{
    constexpr int CH = wg_channels(CI);
    constexpr int NCG = wg_groups(CI);
    constexpr int NRB = wg_row_blocks(CI);
    constexpr int RBW = (NRB + 3) / 4;  // row blocks per wave
    constexpr int CBN = c_blocks(CO);
    constexpr int COP = 16 * CBN;
    constexpr int ROWS = wg_rows(CI);
    constexpr int kXBytes = 3 * CH * kWgXS;
    constexpr int kDyOff = kXBytes;
    constexpr int kBytes = kDyOff + COP * kWgDS;
    static_assert(kBytes <= 160 * 1024, "LDS budget");
    __shared__ __attribute__((aligned(16))) unsigned char smem[kBytes];
    unsigned char* xs = smem;
    unsigned char* dys = smem + kDyOff;

    const int cg = (int)blockIdx.x % NCG;  // channel group: input channels 16 cg .. (CI >= 16)
    const int slot = (int)blockIdx.x / NCG;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int ln = lane % 16, kq = lane / 16;
    const int tiles_x = (m + kTX - 1) / kTX;
    const int tiles_xy = tiles_x * tiles_x;
    const int ntiles = tiles_xy * (m / kTZ);
    const int mh = m / 2;

    // padded output channels stay zero in the dy tile
    for (int i = threadIdx.x; i < (COP - CO) * (kWgDS / 4); i += kBlock)
        reinterpret_cast<unsigned*>(dys + CO * kWgDS)[i] = 0u;

    // the lane's row in each of its row blocks: kind 0 = weight row (A from the
    // shifted halo), 1 = the ones row, 2 = padding; byte offset of its A rows
    int kind[RBW], aoff[RBW], grow[RBW];
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
        const int rb = wave + 4 * j;
        int g = -1, k = 2;
        if (rb < NRB) {
            if constexpr (CI >= 16) {
                if (rb < 27) {
                    g = rb * CI + 16 * cg + ln;
                    k = 0;
                } else if (cg == 0 && ln == 0) {
                    g = 27 * CI;
                    k = 1;
                }
            } else {
                const int gg = 16 * rb + ln;
                if (gg < 27 * CI) {
                    g = gg;
                    k = 0;
                } else if (gg == 27 * CI) {
                    g = gg;
                    k = 1;
                }
            }
        }
        kind[j] = k;
        grow[j] = g;
        aoff[j] = 0;
        if (k == 0) {
            const int t = g / CI, c = (g % CI) - (CI >= 16 ? 16 * cg : 0);
            const int dz = t / 9, dyy = (t / 3) % 3, s = t % 3;  // halo offsets (d + 1)
            aoff[j] = (s * CH + c) * kWgXS + (dz * 10 + dyy) * 16;
        }
    }
    f32x4 acc[RBW][CBN];
#pragma unroll
    for (int j = 0; j < RBW; ++j)
#pragma unroll
        for (int cb = 0; cb < CBN; ++cb)
            acc[j][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 ones = {0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80};

    for (TileWalk walk(ntiles, slot, nslots); walk.t < walk.end; walk.t += walk.step) {
        const int t2 = walk.t;
        const int X0 = (t2 % tiles_x) * kTX;
        const int Y0 = ((t2 / tiles_x) % tiles_x) * kTY;
        const int Z0 = (t2 / tiles_xy) * kTZ;
        __syncthreads();  // the previous tile's reads are done
        // shifted halo copies of the group's CH input channels
        if constexpr (CI >= 16) {
            constexpr int ITEMS = kHalo * 2;  // (voxel, channel octet)
            for (int it = threadIdx.x; it < ITEMS; it += kBlock) {
                const int hv = it >> 1, od = it & 1;
                const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
                const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
                bf16x8 v = zero8();
                if (gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m) {
                    const size_t g = (size_t)(gz * m + gy) * m + gx;
                    v = load8(x + g * CI + 16 * cg + 8 * od);
                    if constexpr (IN == kInUpAdd)
                        v = add8(v, load8(coarse + ((size_t)((gz >> 1) * mh + (gy >> 1)) * mh + (gx >> 1)) * CI +
                                          16 * cg + 8 * od));
                }
#pragma unroll
                for (int sft = 0; sft < 3; ++sft) {
                    const int xl = hx - sft;
                    if (xl >= 0 && xl < 8) {
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                            *reinterpret_cast<unsigned short*>(xs + (sft * CH + 8 * od + c) * kWgXS +
                                                               (hz * 10 + hy) * 16 + 2 * xl) = (unsigned short)v[c];
                    }
                }
            }
        } else {
            for (int hv = threadIdx.x; hv < kHalo; hv += kBlock) {
                const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
                const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
                unsigned u = 0u;
                if (gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m) {
                    const size_t g = (size_t)(gz * m + gy) * m + gx;
                    u = CI == 2 ? reinterpret_cast<const unsigned*>(x)[g] : (unsigned)x[g];
                }
#pragma unroll
                for (int sft = 0; sft < 3; ++sft) {
                    const int xl = hx - sft;
                    if (xl >= 0 && xl < 8) {
#pragma unroll
                        for (int c = 0; c < CI; ++c)
                            *reinterpret_cast<unsigned short*>(xs + (sft * CH + c) * kWgXS + (hz * 10 + hy) * 16 +
                                                               2 * xl) = (unsigned short)(u >> (16 * c));
                    }
                }
            }
        }
        // dy tile, channel-major BF16: dys[co][z * 8 + y][x]
        {
            constexpr int QPV = CO >= 4 ? CO / 4 : 1;  // 4-channel quads per voxel
            constexpr int ITEMS = kTX * kTY * kTZ * QPV;
            for (int it = threadIdx.x; it < ITEMS; it += kBlock) {
                const int vl = it / QPV, qd = it % QPV;
                const int lx = vl % 8, ly = (vl / 8) % 8, lz = vl / 64;
                const int gx = X0 + lx, gy = Y0 + ly, gz = Z0 + lz;
                const bool in = gx < m && gy < m;
                const size_t g = in ? (size_t)(gz * m + gy) * m + gx : 0;
                unsigned char* row = dys + (lz * 8 + ly) * 16 + 2 * lx;
                if constexpr (CO >= 4) {
                    const float4 d = in ? *reinterpret_cast<const float4*>(dy + g * CO + 4 * qd)
                                        : float4{0.f, 0.f, 0.f, 0.f};
                    const unsigned p0 = pack_bf16x2(d.x, d.y), p1 = pack_bf16x2(d.z, d.w);
                    *reinterpret_cast<unsigned short*>(row + (4 * qd + 0) * kWgDS) = (unsigned short)p0;
                    *reinterpret_cast<unsigned short*>(row + (4 * qd + 1) * kWgDS) = (unsigned short)(p0 >> 16);
                    *reinterpret_cast<unsigned short*>(row + (4 * qd + 2) * kWgDS) = (unsigned short)p1;
                    *reinterpret_cast<unsigned short*>(row + (4 * qd + 3) * kWgDS) = (unsigned short)(p1 >> 16);
                } else {
#pragma unroll
                    for (int c = 0; c < CO; ++c) {
                        const float d = in ? dy[g * CO + c] : 0.0f;
                        *reinterpret_cast<unsigned short*>(row + c * kWgDS) = (unsigned short)pack_bf16x2(d, 0.0f);
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const int R = 4 * kb + kq;  // x row of the tile: y = R % 8, z = R / 8
            const int rofs = ((R / 8) * 10 + (R % 8)) * 16;
            mfma_bf16x8 b[CBN];
#pragma unroll
            for (int cb = 0; cb < CBN; ++cb)
                b[cb] = __builtin_bit_cast(mfma_bf16x8,
                                           *reinterpret_cast<const bf16x8*>(dys + (16 * cb + ln) * kWgDS + R * 16));
#pragma unroll
            for (int j = 0; j < RBW; ++j) {
                if (wave + 4 * j >= NRB)
                    continue;  // wave-uniform
                bf16x8 av = zero8();
                if (kind[j] == 0)
                    av = *reinterpret_cast<const bf16x8*>(xs + aoff[j] + rofs);
                else if (kind[j] == 1)
                    av = ones;
                const mfma_bf16x8 a = __builtin_bit_cast(mfma_bf16x8, av);
#pragma unroll
                for (int cb = 0; cb < CBN; ++cb)
                    acc[j][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[cb], acc[j][cb], 0, 0, 0);
            }
        }
    }
    // partial[slot][row][co]: lane holds rows 4 kq + i of each row block, column ln
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
        const int rb = wave + 4 * j;
        if (rb >= NRB)
            continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // the row held in C row 4 kq + i of this block (same mapping as the A rows)
            const int r = 4 * kq + i;
            int g = -1;
            if constexpr (CI >= 16) {
                if (rb < 27)
                    g = rb * CI + 16 * cg + r;
                else if (cg == 0 && r == 0)
                    g = 27 * CI;
            } else {
                g = 16 * rb + r < ROWS ? 16 * rb + r : -1;
            }
            if (g >= 0) {
#pragma unroll
                for (int cb = 0; cb < CBN; ++cb)
                    partial[((size_t)slot * ROWS + g) * COP + 16 * cb + ln] = acc[j][cb][i];
            }
        }
    }
}

// dW[row * CO + co] (rows < 27 * CI) and db[co] (row 27 * CI) = sum over the slots in order
__global__ __launch_bounds__(kBlock) void wgrad_sum_kernel(int rows, int co, int cop, int nslots,
                                                           const float* __restrict__ partial, float* __restrict__ dW,
                                                           float* __restrict__ db)
{
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= rows * co)
        return;
    const int g = idx / co, c = idx % co;
    float s = 0.0f;
    for (int k = 0; k < nslots; ++k)
        s += partial[((size_t)k * rows + g) * cop + c];
    if (g < rows - 1)
        dW[(size_t)g * co + c] = s;
    else
        db[c] = s;
}

template <int CI, int CO>
int launch_wgrad(int m, int in_mode, const unsigned short* x, const unsigned short* coarse, const float* dy,
                 float* dW, float* db, float* partial, hipStream_t s)
{
    const int txy = (m + kTX - 1) / kTX;
    const int tiles = txy * txy * (m / kTZ);
    const int nslots = tiles < kWgMaxSlots ? tiles : kWgMaxSlots;
    const unsigned grid = (unsigned)(nslots * wg_groups(CI));
    if (in_mode == kInUpAdd) {
        if constexpr (CI >= 16)
            conv3_wgrad_mfma_kernel<CI, CO, kInUpAdd><<<grid, kBlock, 0, s>>>(m, x, coarse, dy, nslots, partial);
    } else {
        conv3_wgrad_mfma_kernel<CI, CO, kInPlain><<<grid, kBlock, 0, s>>>(m, x, coarse, dy, nslots, partial);
    }
    WA_LAUNCH_CHECK();
    const int rows = wg_rows(CI);
    wgrad_sum_kernel<<<(unsigned)ceil_div(rows * CO, kBlock), kBlock, 0, s>>>(rows, CO, 16 * c_blocks(CO), nslots,
                                                                           partial, dW, db);
    WA_LAUNCH_CHECK();
    return kOk;
}

// BF16 ReLU mask of an FP32 gradient: out = y > 0 ? dy : 0 (y BF16)
__global__ __launch_bounds__(kBlock) void relu_mask_bf16_kernel(int64_t count, const unsigned short* __restrict__ y,
                                                                const float* __restrict__ dy, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < count)
        out[i] = bf2f(y[i]) > 0.0f ? dy[i] : 0.0f;
}

}  // namespace
}  // namespace wa

namespace {
bool conv_pair_supported(int ci, int co)
{
    static const int pairs[][2] = {{2, 16}, {16, 32}, {32, 64}, {64, 64}, {64, 32}, {32, 16}, {16, 1},
                                   {16, 16}, {1, 16}, {16, 2}};
    for (const auto& p : pairs)
        if (p[0] == ci && p[1] == co)
            return true;
    return false;
}
}  // namespace

// Bytes of the swizzled BF16 weights of a (ci -> co) layer (for one output
// channel: BF16 pairs [27][ci / 2] for the packed-dot kernel).
WA_API size_t warp_amd_conv3_mfma_weight_size(int cin, int cout)
{
    if (cout == 1 && cin >= 16)
        return sizeof(unsigned) * 27 * (size_t)(cin / 2);
    return sizeof(short) * 8 * 64 * (size_t)wa::k_blocks(cin) * wa::c_blocks(cout);
}

// FP32 W[27][cin][cout] -> swizzled BF16 A-operand tiles for a (cin -> cout)
// convolution.  transpose_flip = 1 builds the backward-data operator of a
// (cout -> cin) layer from its weights W[27][cout][cin].
WA_API int warp_amd_conv3_mfma_swizzle(int cin, int cout, int transpose_flip, const float* W, void* wsw, void* stream)
{
    WA_REQUIRE(conv_pair_supported(cin, cout), "conv3_mfma: unsupported channel pair %d -> %d", cin, cout);
    if (cout == 1 && cin >= 16) {
        const int pairs = 27 * (cin / 2);
        wa::pack_co1_kernel<<<(unsigned)wa::ceil_div(pairs, 256), 256, 0, wa::as_stream(stream)>>>(
            cin, transpose_flip, W, static_cast<unsigned*>(wsw)
        );
        WA_LAUNCH_CHECK();
        return wa::kOk;
    }
    const int total = (int)(warp_amd_conv3_mfma_weight_size(cin, cout) / 16);
    wa::swizzle_kernel<<<(unsigned)wa::ceil_div(total, 256), 256, 0, wa::as_stream(stream)>>>(
        cin, cout, transpose_flip, W, static_cast<wa::bf16x8*>(wsw), total
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// BF16 3x3x3 convolution on an m^3 grid (m a multiple of 4; even with
// pooling or up-add).  in_mode 0: input x; 1: x + upsample(coarse).  flags:
// 1 ReLU, 2 also write the 2x average pool, 4 FP32 output yf (else BF16 y),
// 8 accumulate into yf.
WA_API int warp_amd_conv3_mfma(int m, int cin, int cout, int in_mode, const void* x, const void* coarse,
                               const void* wsw, const float* bias, int flags, void* y, float* yf, void* pooled,
                               void* stream)
{
    WA_REQUIRE(conv_pair_supported(cin, cout), "conv3_mfma: unsupported channel pair %d -> %d", cin, cout);
    WA_REQUIRE(m >= 4 && m % 4 == 0, "conv3_mfma: grid side %d must be a positive multiple of 4", m);
    WA_REQUIRE(in_mode == 0 || (in_mode == 1 && cin > 2 && coarse), "conv3_mfma: bad input mode");
    WA_REQUIRE(!(flags & wa::kOutPool) || (pooled && cout % 16 == 0 && !(flags & wa::kOutF32)),
               "conv3_mfma: pooling needs a BF16 output with a multiple of 16 channels");
    WA_REQUIRE((flags & wa::kOutF32) ? yf != nullptr : (y != nullptr && cout % 16 == 0),
               "conv3_mfma: missing output (BF16 output needs a multiple of 16 channels)");
    const unsigned short* xi = static_cast<const unsigned short*>(x);
    const unsigned short* ci = static_cast<const unsigned short*>(coarse);
    unsigned short* yo = static_cast<unsigned short*>(y);
    unsigned short* po = static_cast<unsigned short*>(pooled);
    hipStream_t s = wa::as_stream(stream);
#define WA_CONV_CASE(A, B)                                                                                       \
    if (cin == A && cout == B)                                                                                   \
        return wa::launch_conv<A, B>(m, in_mode, xi, ci, wsw, bias, flags, yo, yf, po, s);
    WA_CONV_CASE(2, 16)
    WA_CONV_CASE(16, 32)
    WA_CONV_CASE(32, 64)
    WA_CONV_CASE(64, 64)
    WA_CONV_CASE(64, 32)
    WA_CONV_CASE(32, 16)
    if (cin == 16 && cout == 1) {
        WA_REQUIRE(flags & wa::kOutF32, "conv3_mfma: a one-channel output is FP32");
        return wa::launch_co1<16>(m, in_mode, xi, ci, wsw, bias, flags, yf, s);
    }
    WA_CONV_CASE(16, 16)
    WA_CONV_CASE(1, 16)
    WA_CONV_CASE(16, 2)
#undef WA_CONV_CASE
    return wa::kErrUnsupported;
}

WA_API int warp_amd_f32_to_bf16(int64_t count, const float* in, void* out, void* stream)
{
    if (count <= 0)
        return wa::kOk;
    wa::f32_to_bf16_kernel<<<(unsigned)wa::ceil_div(count, 256), 256, 0, wa::as_stream(stream)>>>(
        count, in, static_cast<unsigned short*>(out)
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_bf16_to_f32(int64_t count, const void* in, float* out, void* stream)
{
    if (count <= 0)
        return wa::kOk;
    wa::bf16_to_f32_kernel<<<(unsigned)wa::ceil_div(count, 256), 256, 0, wa::as_stream(stream)>>>(
        count, static_cast<const unsigned short*>(in), out
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

// BF16 [nvox][2] network input (r * inv_scale, logk)
WA_API int warp_amd_unet_pack_input_bf16(int64_t nvox, const double* r, const float* logk_std,
                                         const double* inv_scale, void* out, void* stream)
{
    const bool quads = nvox % 4 == 0 && ((uintptr_t)r | (uintptr_t)logk_std | (uintptr_t)out) % 16 == 0;
    if (quads)
        wa::pack_input_bf16_x4_kernel<<<(unsigned)wa::ceil_div(nvox / 4, 256), 256, 0, wa::as_stream(stream)>>>(
            nvox / 4, r, logk_std, inv_scale, static_cast<unsigned short*>(out)
        );
    else
        wa::pack_input_bf16_kernel<<<(unsigned)wa::ceil_div(nvox, 256), 256, 0, wa::as_stream(stream)>>>(
            nvox, r, logk_std, inv_scale, static_cast<unsigned short*>(out)
        );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

namespace {
bool wgrad_pair_supported(int ci, int co)
{
    static const int pairs[][2] = {{2, 16}, {16, 32}, {32, 64}, {64, 64}, {64, 32}, {32, 16}, {16, 1}};
    for (const auto& p : pairs)
        if (p[0] == ci && p[1] == co)
            return true;
    return false;
}
}  // namespace

// Bytes of workspace of warp_amd_conv3_mfma_wgrad on an m^3 grid.
WA_API size_t warp_amd_conv3_mfma_wgrad_workspace_size(int m, int cin, int cout)
{
    if (m <= 0)
        return 0;
    const int64_t txy = wa::ceil_div(m, wa::kTX);
    const int64_t tiles = txy * txy * (m / wa::kTZ);
    const int64_t slots = tiles < wa::kWgMaxSlots ? tiles : wa::kWgMaxSlots;
    return sizeof(float) * (size_t)slots * wa::wg_rows(cin) * 16 * wa::c_blocks(cout);
}

// Weight and bias gradients of a BF16 (cin -> cout) 3x3x3 convolution on an
// m^3 grid (m a multiple of 4): dW[27][cin][cout] = sum_v x~[v + d_t][ci]
// dy[v][co], db[cout] = sum_v dy[v][co], both overwritten.  x: BF16 layer
// input [m^3][cin] (in_mode 1: x + upsample(coarse), as the forward formed
// it); dy: FP32 [m^3][cout], rounded to BF16; FP32 accumulation, partial
// sums per workgroup added in a fixed order (deterministic).
WA_API int warp_amd_conv3_mfma_wgrad(int m, int cin, int cout, int in_mode, const void* x, const void* coarse,
                                     const float* dy, float* dW, float* db, void* workspace, size_t workspace_bytes,
                                     void* stream)
{
    WA_REQUIRE(wgrad_pair_supported(cin, cout), "conv3_mfma_wgrad: unsupported channel pair %d -> %d", cin, cout);
    WA_REQUIRE(m >= 4 && m % 4 == 0, "conv3_mfma_wgrad: grid side %d must be a positive multiple of 4", m);
    WA_REQUIRE(in_mode == 0 || (in_mode == 1 && cin >= 16 && coarse), "conv3_mfma_wgrad: bad input mode");
    WA_REQUIRE(x && dy && dW && db, "conv3_mfma_wgrad: null array");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_conv3_mfma_wgrad_workspace_size(m, cin, cout),
               "conv3_mfma_wgrad: workspace too small");
    const unsigned short* xi = static_cast<const unsigned short*>(x);
    const unsigned short* ci = static_cast<const unsigned short*>(coarse);
    float* pw = static_cast<float*>(workspace);
    hipStream_t s = wa::as_stream(stream);
#define WA_WGRAD_CASE(A, B)                                                                                      \
    if (cin == A && cout == B)                                                                                   \
        return wa::launch_wgrad<A, B>(m, in_mode, xi, ci, dy, dW, db, pw, s);
    WA_WGRAD_CASE(2, 16)
    WA_WGRAD_CASE(16, 32)
    WA_WGRAD_CASE(32, 64)
    WA_WGRAD_CASE(64, 64)
    WA_WGRAD_CASE(64, 32)
    WA_WGRAD_CASE(32, 16)
    WA_WGRAD_CASE(16, 1)
#undef WA_WGRAD_CASE
    return wa::kErrUnsupported;
}

// out = y > 0 ? dy : 0 with y BF16 (the ReLU of a BF16 forward), dy and out FP32
WA_API int warp_amd_relu_mask_bf16(int64_t count, const void* y, const float* dy, float* out, void* stream)
{
    if (count <= 0)
        return wa::kOk;
    wa::relu_mask_bf16_kernel<<<(unsigned)wa::ceil_div(count, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        count, static_cast<const unsigned short*>(y), dy, out
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}
