// BF16 MFMA implicit-GEMM 3x3x3 convolution for the learned preconditioner.
//
// No reference counterpart.  Same conventions as the FP32 kernels (unet.hip):
// channels-last fields [voxel][channel] on an m^3 grid, voxel = x + m (y + m z),
// weights W[tap][ci][co], tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding.
// Activations and weights are BF16 (round to nearest even), products are
// accumulated in FP32 by v_mfma_f32_16x16x16_bf16:
//
//   C[co][voxel] = sum_k A[co][k] B[k][voxel],   k = (tap, ci)
//
// Operand lane layout of the 16x16x16 instruction (64 lanes, l = 16 q + n):
// A: lane holds row n, k = 4q..4q+3;  B: lane holds column n, k = 4q..4q+3;
// C: lane holds column n, rows 4q..4q+3.  With A = weights and B = input,
// a lane's B operand is 4 consecutive channels of one voxel (one 8-byte
// load) and its C fragment is 4 consecutive output channels of one voxel
// (one 8-byte store).  Weights are pre-swizzled into A-operand order so each
// lane reads its 4 values with one 8-byte load.
//
// Work split: a wavefront owns a 4x4x4 voxel cube; column n = (y%4) * 4 +
// (x%4) and the four MFMA column tiles g = 0..3 are the cube's z slices.  The
// 2x2x2 average pool of the output is then a sum over tile pairs (g, g+1)
// plus the lane pairs n^1 (x) and n^4 (y), done in registers.  Four cubes
// share one workgroup and one LDS copy of their input halo.
//
// Input channel counts 1 and 2 (first layer, backward of the last) fold the
// taps into K: k = tap * CI + ci, padded with zero weights to 16 * KB.
#include "common.h"

namespace wa {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
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

__host__ __device__ constexpr int k_blocks(int ci) { return ci <= 2 ? (27 * ci + 15) / 16 : ci / 16; }
__host__ __device__ constexpr int tap_loops(int ci) { return ci <= 2 ? 1 : 27; }
__host__ __device__ constexpr int c_blocks(int co) { return (co + 15) / 16; }

// A-operand tiles: Wsw[((t * KB + kb) * CB + cb) * 64 + lane] = 4 BF16 of
// W(tap, ci, co) for co = 16 cb + lane % 16, k = 16 kb + 4 (lane / 16) + j.
// transpose_flip: use W'(tap, ci, co) = W[26 - tap][co][ci] with W stored
// for a (co -> ci) layer, i.e. the backward-data convolution.
__global__ void swizzle_kernel(int ci, int co, int transpose_flip, const float* __restrict__ W, bf16x4* __restrict__ out,
                               int total)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total)
        return;
    const int KB = k_blocks(ci), CB = c_blocks(co);
    const int lane = idx % 64;
    const int cb = (idx / 64) % CB;
    const int kb = (idx / (64 * CB)) % KB;
    const int t0 = idx / (64 * CB * KB);
    const int o = cb * 16 + lane % 16;
    bf16x4 v;
    for (int j = 0; j < 4; ++j) {
        const int k = kb * 16 + 4 * (lane / 16) + j;
        int tap, c;
        if (ci <= 2) {
            tap = k / ci;
            c = k % ci;
        } else {
            tap = t0;
            c = k;
        }
        float w = 0.0f;
        if (o < co && tap < 27 && c < ci)
            w = transpose_flip ? W[((size_t)(26 - tap) * co + o) * ci + c] : W[((size_t)tap * ci + c) * co + o];
        v[j] = (short)f2bf(w);
    }
    out[idx] = v;
}

__device__ __forceinline__ bf16x4 zero4()
{
    bf16x4 z = {0, 0, 0, 0};
    return z;
}

__device__ __forceinline__ bf16x4 load4(const unsigned short* p) { return *reinterpret_cast<const bf16x4*>(p); }

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    return z;
}

__device__ __forceinline__ bf16x8 load8(const unsigned short* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ bf16x8 add8(bf16x8 a, bf16x8 b)
{
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        r[j] = (short)f2bf(bf2f((unsigned short)a[j]) + bf2f((unsigned short)b[j]));
    return r;
}

__device__ __forceinline__ bf16x4 add4(bf16x4 a, bf16x4 b)
{
    bf16x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        r[j] = (short)f2bf(bf2f((unsigned short)a[j]) + bf2f((unsigned short)b[j]));
    return r;
}

enum InMode { kInPlain = 0, kInUpAdd = 1 };
enum OutFlags { kOutRelu = 1, kOutPool = 2, kOutF32 = 4, kOutAccumulate = 8 };

// Workgroup tile: 8 x 8 x 4 voxels = four 4x4x4 cubes (wave w: cube (w & 1,
// w >> 1) in x, y).  Its 10 x 10 x 6 halo is staged once in LDS (the
// upsample + skip add applied there, once per voxel), then every wave reads
// its B operands from LDS for all 27 taps.
// Persistent tile loops: workgroup b is dispatched to XCD b % kXcds (the
// MI355X command processor deals workgroups to its 8 XCDs round-robin), so
// each XCD walks its own contiguous chunk of tiles: tiles that share halo
// voxels then run on the same XCD and meet in its L2.  Any other mapping
// only costs locality, never correctness.
constexpr int kXcds = 8;

struct TileWalk {
    int t, end, step;
    __device__ __forceinline__ explicit TileWalk(int ntiles)
    {
        if (gridDim.x >= kXcds && gridDim.x % kXcds == 0) {
            const int xcd = blockIdx.x % kXcds;
            const int chunk = (ntiles + kXcds - 1) / kXcds;
            t = xcd * chunk + blockIdx.x / kXcds;
            end = min(ntiles, (xcd + 1) * chunk);
            step = gridDim.x / kXcds;
        } else {
            t = blockIdx.x;
            end = ntiles;
            step = gridDim.x;
        }
    }
};

constexpr int kTX = 8, kTY = 8, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;

// LDS words (4 bytes) per halo voxel: CI channels of BF16, plus 4 words of
// padding for CI >= 16 to spread the 8-byte reads over the banks (a
// multiple of 4 words keeps the 16-byte staging stores aligned).
__host__ __device__ constexpr int lds_stride_words(int ci) { return ci <= 2 ? 1 : ci / 2 + 4; }

// Stage the 10 x 10 x 6 input halo of the tile at (X0, Y0, Z0) in LDS
// (zeros outside the grid; with kInUpAdd the input is x + upsample(coarse),
// rounded to BF16 once per voxel).
template <int CI, int IN>
__device__ __forceinline__ void stage_tile(unsigned* tile, int m, int X0, int Y0, int Z0,
                                           const unsigned short* __restrict__ x,
                                           const unsigned short* __restrict__ coarse)
{
    constexpr int SW = lds_stride_words(CI);
    const int mh = m / 2;
    // All loads of a batch are issued before any is used: the halo is ~10
    // loads per thread, and issuing them one round trip at a time made the
    // staging, not the MFMAs, the cost of the layer.
    if constexpr (CI <= 2) {
        constexpr int ITERS = (kHalo + kBlock - 1) / kBlock;
        unsigned v[ITERS];
#pragma unroll
        for (int u = 0; u < ITERS; ++u) {
            const int hv = u * kBlock + threadIdx.x;
            const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
            const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
            v[u] = 0;
            if (hv < kHalo && gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m) {
                const size_t g = (size_t)(gz * m + gy) * m + gx;
                v[u] = CI == 2 ? reinterpret_cast<const unsigned*>(x)[g] : (unsigned)x[g];
            }
        }
#pragma unroll
        for (int u = 0; u < ITERS; ++u) {
            const int hv = u * kBlock + threadIdx.x;
            if (hv < kHalo)
                tile[hv] = v[u];
        }
    } else {
        constexpr int QPV = CI / 8;  // 16-byte channel octets per voxel
        constexpr int TOTAL = kHalo * QPV;
        constexpr int ITERS = (TOTAL + kBlock - 1) / kBlock;
        constexpr int BATCH = ITERS < 4 ? ITERS : 4;
        for (int base = 0; base < ITERS; base += BATCH) {
            bf16x8 v[BATCH];
            bf16x8 cv[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int c = (base + u) * kBlock + threadIdx.x;
                const int hv = c / QPV, od = c % QPV;
                const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
                const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
                v[u] = zero8();
                cv[u] = zero8();
                if (base + u < ITERS && c < TOTAL && gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m) {
                    v[u] = load8(x + ((size_t)(gz * m + gy) * m + gx) * CI + 8 * od);
                    if constexpr (IN == kInUpAdd)
                        cv[u] = load8(coarse + ((size_t)((gz >> 1) * mh + (gy >> 1)) * mh + (gx >> 1)) * CI + 8 * od);
                }
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int c = (base + u) * kBlock + threadIdx.x;
                if (base + u < ITERS && c < TOTAL) {
                    bf16x8 w = v[u];
                    if constexpr (IN == kInUpAdd)
                        w = add8(w, cv[u]);
                    *reinterpret_cast<bf16x8*>(&tile[(c / QPV) * SW + 4 * (c % QPV)]) = w;
                }
            }
        }
    }
}

// x: input [m^3][CI] BF16; with kInUpAdd the input is x[v] + coarse[v / 2]
// (coarse: [(m/2)^3][CI]), rounded to BF16.  Output: y BF16 [m^3][CO] or, with
// kOutF32, yf FP32 [m^3][CO] (kOutAccumulate adds to yf); pooled BF16
// [(m/2)^3][CO] with kOutPool.
// CBG of the CO / 16 output-channel blocks per workgroup (blockIdx.z picks
// which), so that coarse grids still launch enough workgroups.
template <int CI, int CO, int IN, int CBG>
__global__ __launch_bounds__(kBlock) void conv3_mfma_kernel(
    int m, const unsigned short* __restrict__ x, const unsigned short* __restrict__ coarse,
    const bf16x4* __restrict__ wsw, const float* __restrict__ bias, int flags, unsigned short* __restrict__ y,
    float* __restrict__ yf, unsigned short* __restrict__ pooled
)
{
    constexpr int KB = k_blocks(CI);
    constexpr int CB = c_blocks(CO);
    constexpr int TL = tap_loops(CI);
    constexpr int SW = lds_stride_words(CI);
    static_assert(CB % CBG == 0, "channel blocks per workgroup must divide the channel blocks");
    __shared__ unsigned tile[kHalo * SW];

    const int tiles_x = (m + kTX - 1) / kTX;
    const int tiles_xy = tiles_x * tiles_x;
    const int nsplit = CB / CBG;
    const int ntiles = tiles_xy * (m / kTZ) * nsplit;
    const int mh = m / 2;
    // persistent loop over tiles; the channel split varies fastest so that
    // concurrent workgroups share one input tile in L2
    for (TileWalk walk(ntiles); walk.t < walk.end; walk.t += walk.step) {
    const int tile_id = walk.t;
    const int cb0 = (tile_id % nsplit) * CBG;
    const int t2 = tile_id / nsplit;
    const int X0 = (t2 % tiles_x) * kTX;
    const int Y0 = ((t2 / tiles_x) % tiles_x) * kTY;
    const int Z0 = (t2 / tiles_xy) * kTZ;

    __syncthreads();  // the previous tile's LDS reads are done
    stage_tile<CI, IN>(tile, m, X0, Y0, Z0, x, coarse);
    __syncthreads();

    const int lane = threadIdx.x % 64;
    const int w = threadIdx.x / 64;
    const int n = lane % 16, q = lane / 16;
    const int lx = 4 * (w & 1) + (n & 3);  // position in the tile
    const int ly = 4 * (w >> 1) + (n >> 2);
    const int vx = X0 + lx, vy = Y0 + ly, vz0 = Z0;
    const bool valid = (X0 + 4 * (w & 1)) < m && (Y0 + 4 * (w >> 1)) < m;  // whole cube in range

    f32x4 acc[CBG][4];
#pragma unroll
    for (int cb = 0; cb < CBG; ++cb)
#pragma unroll
        for (int g = 0; g < 4; ++g)
            acc[cb][g] = f32x4{0.f, 0.f, 0.f, 0.f};
    // CI <= 2: the halo offset of every tap this lane reads (-1: padding tap >= 27)
    constexpr int NT = CI == 2 ? 2 : 4;  // taps per lane and k-block
    int toff[CI <= 2 ? KB : 1][NT];
    if constexpr (CI <= 2) {
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const int tap = CI == 2 ? kb * 8 + 2 * q + i : kb * 16 + 4 * q + i;
                toff[kb][i] = tap < 27 ? ((tap / 9) * kHY + ly + (tap / 3) % 3) * kHX + lx + tap % 3 : -1;
            }
    }

    for (int t = 0; t < TL; ++t) {
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
        // halo coordinates of this lane's neighbour for slice g: (lx+1+dx, ly+1+dy, g+1+dz)
        const int hbase = ((1 + dz) * kHY + (ly + 1 + dy)) * kHX + (lx + 1 + dx);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            bf16x4 a[CBG];
#pragma unroll
            for (int cb = 0; cb < CBG; ++cb)
                a[cb] = wsw[((t * KB + kb) * CB + cb0 + cb) * 64 + lane];
            bf16x4 b[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                if constexpr (CI <= 2) {
                    // k = 16 kb + 4 q + j -> (tap, channel) = (k / CI, k % CI); halo
                    // offsets of the taps precomputed in toff
                    bf16x4 v;
                    const int gz = g * (kHX * kHY);
                    if constexpr (CI == 2) {
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            const unsigned u = toff[kb][i] >= 0 ? tile[gz + toff[kb][i]] : 0u;
                            v[2 * i] = (short)(u & 0xffffu);  // channels 0, 1 of one tap
                            v[2 * i + 1] = (short)(u >> 16);
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const unsigned u = toff[kb][i] >= 0 ? tile[gz + toff[kb][i]] : 0u;
                            v[i] = (short)(u & 0xffffu);
                        }
                    }
                    b[g] = v;
                } else {
                    const int hv = hbase + g * (kHX * kHY);
                    b[g] = *reinterpret_cast<const bf16x4*>(&tile[hv * SW + 8 * kb + 2 * q]);
                }
            }
#pragma unroll
            for (int cb = 0; cb < CBG; ++cb)
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    acc[cb][g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[cb], b[g], acc[cb][g], 0, 0, 0);
        }
    }
    if (!valid)
        continue;  // whole wavefront (only when m < 8); still joins the next tile's barriers

    // epilogue: lane holds output channels 16 cb + 4 q + i of voxel (vx, vy, vz0 + g)
    const bool relu = flags & kOutRelu;
#pragma unroll
    for (int cb = 0; cb < CBG; ++cb) {
        const int o0 = (cb0 + cb) * 16 + 4 * q;
        float bsv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            bsv[i] = (bias && o0 + i < CO) ? bias[o0 + i] : 0.0f;
        float psum[2][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const size_t v = (size_t)((vz0 + g) * m + vy) * m + vx;
            float val[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                val[i] = acc[cb][g][i] + bsv[i];
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
                bf16x4 st;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    st[i] = (short)f2bf(val[i]);
                    val[i] = bf2f((unsigned short)st[i]);  // pool what was stored
                }
                if constexpr (CO % 16 == 0)
                    *reinterpret_cast<bf16x4*>(y + v * CO + o0) = st;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                psum[g / 2][i] = (g % 2) ? psum[g / 2][i] + val[i] : val[i];
        }
        if (flags & kOutPool) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                bf16x4 st;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float sv = psum[h][i];
                    sv += __shfl_xor(sv, 1);
                    sv += __shfl_xor(sv, 4);
                    st[i] = (short)f2bf(sv * 0.125f);
                }
                if ((n & 1) == 0 && (n & 4) == 0) {
                    const size_t pv = (size_t)(((vz0 >> 1) + h) * mh + (vy >> 1)) * mh + (vx >> 1);
                    if constexpr (CO % 16 == 0)
                        *reinterpret_cast<bf16x4*>(pooled + pv * CO + o0) = st;
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

// Halo (HX x HY x HZ voxels) of one tile held in registers between its loads
// and its LDS stores (CI >= 16: 16-byte channel octets, CI / 8 per voxel).
template <int CI, int IN, int HX = kHX, int HY = kHY, int HZ = kHZ> struct HaloRegs {
    static constexpr int QPV = CI / 8;
    static constexpr int TOTAL = HX * HY * HZ * QPV;
    static constexpr int ITERS = (TOTAL + kBlock - 1) / kBlock;
    bf16x8 v[ITERS];
    bf16x8 cv[IN == kInUpAdd ? ITERS : 1];

    __device__ __forceinline__ void load(int m, int X0, int Y0, int Z0, const unsigned short* __restrict__ x,
                                         const unsigned short* __restrict__ coarse)
    {
        const int mh = m / 2;
#pragma unroll
        for (int u = 0; u < ITERS; ++u) {
            const int c = u * kBlock + threadIdx.x;
            const int hv = c / QPV, od = c % QPV;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
            v[u] = zero8();
            if constexpr (IN == kInUpAdd)
                cv[u] = zero8();
            if (c < TOTAL && gx >= 0 && gx < m && gy >= 0 && gy < m && gz >= 0 && gz < m) {
                v[u] = load8(x + ((size_t)(gz * m + gy) * m + gx) * CI + 8 * od);
                if constexpr (IN == kInUpAdd)
                    cv[u] = load8(coarse + ((size_t)((gz >> 1) * mh + (gy >> 1)) * mh + (gx >> 1)) * CI + 8 * od);
            }
        }
    }

    __device__ __forceinline__ void store(unsigned* tile) const
    {
        constexpr int SW = lds_stride_words(CI);
#pragma unroll
        for (int u = 0; u < ITERS; ++u) {
            const int c = u * kBlock + threadIdx.x;
            if (c < TOTAL) {
                bf16x8 w = v[u];
                if constexpr (IN == kInUpAdd)
                    w = add8(w, cv[u]);
                *reinterpret_cast<bf16x8*>(&tile[(c / QPV) * SW + 4 * (c % QPV)]) = w;
            }
        }
    }
};

// One output channel (the network's last layer): MFMA would waste 15 of its
// 16 rows, so each thread computes voxels with packed BF16 dot products,
// v_dot2_f32_bf16, against weights wpk[tap][ci / 2] (BF16 pairs), reading the
// LDS-staged halo; every voxel adds its taps in the order t = 0..26.  Output
// FP32; flags: ReLU, accumulate.  Tiles are 8 x 8 x 8 with two z-adjacent
// voxels per thread: their 3 x 3 x 4 halo voxels are read from LDS once for
// both (36 reads instead of 54) and a tile's halo is 1.95x its voxels (2.34x
// for 8 x 8 x 4 tiles of one voxel per thread: 64.0 us against 62.6 at
// 128^3).  The halo of the workgroup's next tile is loaded into registers
// while the dot products of the current one run.
constexpr int kTZ2 = 8;
constexpr int kHZ2 = kTZ2 + 2;

template <int CI, int IN>
__global__ __launch_bounds__(kBlock) void conv3_co1_z2_kernel(int m, const unsigned short* __restrict__ x,
                                                             const unsigned short* __restrict__ coarse,
                                                             const unsigned* __restrict__ wpk,
                                                             const float* __restrict__ bias, int flags,
                                                             float* __restrict__ yf)
{
    constexpr int SW = lds_stride_words(CI);
    static_assert(CI >= 16 && CI % 16 == 0, "one-output kernel expects a multiple of 16 input channels");
    __shared__ unsigned tile[kHX * kHY * kHZ2 * SW];
    const int tiles_x = (m + kTX - 1) / kTX;
    const int tiles_z = (m + kTZ2 - 1) / kTZ2;
    const int ntiles = tiles_x * tiles_x * tiles_z;
    TileWalk walk(ntiles);
    auto origin = [&](int t, int& X0, int& Y0, int& Z0) {
        X0 = (t % tiles_x) * kTX;
        Y0 = ((t / tiles_x) % tiles_x) * kTY;
        Z0 = (t / (tiles_x * tiles_x)) * kTZ2;
    };
    HaloRegs<CI, IN, kHX, kHY, kHZ2> halo;
    if (walk.t < walk.end) {
        int X0, Y0, Z0;
        origin(walk.t, X0, Y0, Z0);
        halo.load(m, X0, Y0, Z0, x, coarse);
    }
    const float b0 = bias ? bias[0] : 0.0f;
    const int lx = threadIdx.x & 7, ly = (threadIdx.x >> 3) & 7, lz = threadIdx.x >> 6;  // 8 x 8 x 4 pairs
    for (; walk.t < walk.end; walk.t += walk.step) {
    int X0, Y0, Z0;
    origin(walk.t, X0, Y0, Z0);
    __syncthreads();  // the previous tile's LDS reads are done
    halo.store(tile);
    __syncthreads();
    if (walk.t + walk.step < walk.end) {
        int X1, Y1, Z1;
        origin(walk.t + walk.step, X1, Y1, Z1);
        halo.load(m, X1, Y1, Z1, x, coarse);
    }
    float acc0 = b0, acc1 = b0;
#pragma unroll 1
    for (int dz = 0; dz < 4; ++dz) {
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int hv = ((2 * lz + dz) * kHY + (ly + dy)) * kHX + (lx + dx);
                const unsigned* src = &tile[hv * SW];
                unsigned v[CI / 2];
#pragma unroll
                for (int j = 0; j < CI / 2; ++j)
                    v[j] = src[j];
                if (dz < 3) {
                    const unsigned* w = wpk + ((dz * 3 + dy) * 3 + dx) * (CI / 2);
#pragma unroll
                    for (int j = 0; j < CI / 2; ++j)
                        acc0 = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, v[j]),
                                                               __builtin_bit_cast(bf16x2, w[j]), acc0, false);
                }
                if (dz > 0) {
                    const unsigned* w = wpk + (((dz - 1) * 3 + dy) * 3 + dx) * (CI / 2);
#pragma unroll
                    for (int j = 0; j < CI / 2; ++j)
                        acc1 = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, v[j]),
                                                               __builtin_bit_cast(bf16x2, w[j]), acc1, false);
                }
            }
    }
    const int vx = X0 + lx, vy = Y0 + ly;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int vz = Z0 + 2 * lz + h;
        float acc = h ? acc1 : acc0;
        if (vx < m && vy < m && vz < m) {
            if (flags & kOutRelu)
                acc = fmaxf(acc, 0.0f);
            float* d = yf + (size_t)(vz * m + vy) * m + vx;
            *d = (flags & kOutAccumulate) ? *d + acc : acc;
        }
    }
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
    const int txy = (m + kTX - 1) / kTX;
    const unsigned* w = static_cast<const unsigned*>(wpk);
    const int ntiles = txy * txy * ((m + kTZ2 - 1) / kTZ2);
    if (in_mode == kInUpAdd)
        launch_tiles<conv3_co1_z2_kernel<CI, kInUpAdd>>(ntiles, s, m, x, coarse, w, bias, flags, yf);
    else
        launch_tiles<conv3_co1_z2_kernel<CI, kInPlain>>(ntiles, s, m, x, coarse, w, bias, flags, yf);
    WA_LAUNCH_CHECK();
    return kOk;
}

template <int CI, int CO>
int launch_conv(int m, int in_mode, const unsigned short* x, const unsigned short* coarse, const void* wsw,
                const float* bias, int flags, unsigned short* y, float* yf, unsigned short* pooled, hipStream_t s)
{
    constexpr int CB = c_blocks(CO);
    const int txy = (m + kTX - 1) / kTX;
    const int tiles = txy * txy * (m / kTZ);
    // one channel block per workgroup while that is needed to fill the chip
    const bool split = CB > 1 && tiles < 1024;
    const int ntiles = tiles * (split ? CB : 1);
    const bf16x4* w = static_cast<const bf16x4*>(wsw);
    if (in_mode == kInUpAdd) {
        if constexpr (CI > 2) {
            if (split)
                launch_tiles<conv3_mfma_kernel<CI, CO, kInUpAdd, 1>>(ntiles, s, m, x, coarse, w, bias, flags, y, yf,
                             pooled);
            else
                launch_tiles<conv3_mfma_kernel<CI, CO, kInUpAdd, CB>>(ntiles, s, m, x, coarse, w, bias, flags, y, yf,
                             pooled);
        }
    } else {
        if (split)
            launch_tiles<conv3_mfma_kernel<CI, CO, kInPlain, 1>>(ntiles, s, m, x, coarse, w, bias, flags, y, yf, pooled);
        else
            launch_tiles<conv3_mfma_kernel<CI, CO, kInPlain, CB>>(ntiles, s, m, x, coarse, w, bias, flags, y, yf,
                         pooled);
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
    return sizeof(short) * 4 * 64 * (size_t)wa::tap_loops(cin) * wa::k_blocks(cin) * wa::c_blocks(cout);
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
    const int total = (int)(warp_amd_conv3_mfma_weight_size(cin, cout) / 8);
    wa::swizzle_kernel<<<(unsigned)wa::ceil_div(total, 256), 256, 0, wa::as_stream(stream)>>>(
        cin, cout, transpose_flip, W, static_cast<wa::bf16x4*>(wsw), total
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
    wa::pack_input_bf16_kernel<<<(unsigned)wa::ceil_div(nvox, 256), 256, 0, wa::as_stream(stream)>>>(
        nvox, r, logk_std, inv_scale, static_cast<unsigned short*>(out)
    );
    WA_LAUNCH_CHECK();
    return wa::kOk;
}
