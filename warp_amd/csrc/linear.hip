// Conjugate gradient on gfx950.
//
// Replaces the reference CG iteration (warp/_src/optim/linear.py:893-1121)
// and its JIT kernels: TiledDot (linear.py:309-612), _initialize_tolerance
// (2051-2083), _diag_mv_kernel (2271-2287), _cg_kernel_1/2 (2332-2373) and
// the host loop _run_solver_loop (2182-2243).
//
// Precision policy: vectors stay in the system scalar type S (fp32 or fp64);
// every dot product is accumulated in FP64 with a deterministic two-level
// reduction, and alpha/beta are formed in FP64.
//
// One iteration = three fused kernels on a device-resident state and no
// one-workgroup finalize kernel: each reduction leaves one FP64 partial per
// workgroup, and every workgroup of the next kernel sums all of them in the
// same fixed order in its prologue (bit-identical in every workgroup; the
// grids are capped so that these sums stay a few KB of L2 reads):
//   spmv     : Ap = A p, partials of p.Ap
//   update   : alpha = rz / pAp (0 once converged); x += alpha p,
//              r -= alpha Ap, z = M r, partials of r.r and r.z
//   direction: beta = rz_new / rz_old, done; p = z + beta p
// A field of the state is never written by a kernel whose workgroups read
// it: workgroup 0 of each kernel records what the next one needs (r.z of the
// previous iterate, `done` as the update saw it, the iteration count).
// Small systems (launch-bound: every kernel costs a few microseconds however
// little it does) fold the direction pass into the next SpMV: its prologue
// sums the update's partials into beta, and it evaluates p' = z + beta p
// entry by entry where the product gathers it (the same fma, so the iterates
// are bit-identical) and writes p' to the other of two direction buffers,
// since workgroups still gather p while others write p'.  r.z of the
// previous iterate is +inf after (re)initialisation, which makes beta 0 and
// p' = z.  Large systems keep the separate pass: gathering two vectors costs
// the HBM-bound SpMV more than the pass it saves.
// The convergence flag `done` (r.r <= atol^2) makes later kernels no-ops, the
// device-side equivalent of the reference's frozen alpha/beta.
#include "bsr_core.h"

#include <cmath>
#include <vector>

namespace wa {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxGrid = 2048;  // generic vector kernels

struct CgState {
    double bb;
    double rr;       // r.r of the current iterate (host-visible)
    double rz;       // r.z of the current iterate
    double rz_prev;  // r.z of the previous iterate; +inf after (re)initialisation, so beta = 0
    double pAp;
    double atol_sq;
    double alpha;
    double beta;
    double rtol;
    double atol;
    int iter;       // iterations done (host-visible)
    int iter_copy;  // the SpMV's copy of iter (fused direction buffers), read by the update
    int done;       // r.r <= atol^2 for the current iterate
    int done_prev;  // `done` as the update saw it, read by the direction pass
};

enum PrecKind { kPrecNone = 0, kPrecDiag = 1, kPrecBlock = 2 };

// Grid caps.  Every reduction of an iteration leaves one FP64 partial per
// workgroup, and the next kernel's workgroups each sum all of them in the
// same fixed order in their prologue (so no one-workgroup finalize kernel
// sits between two work kernels): the SpMV walks its row tiles with at most
// kMaxSpmvGrid workgroups, the update and direction passes grid-stride with
// at most kMaxVecGrid.
constexpr int kMaxSpmvGrid = 2048;
constexpr int kMaxVecGrid = 1024;

struct CgWs {
    CgState* state;
    double* init_partials;  // 3 * max_init_blocks (b.b, r.r, r.z of the initial residual)
    double* spmv_partials;  // kMaxSpmvGrid (p.Ap)
    double* upd_partials;   // 2 * kMaxVecGrid (r.r, r.z)
    int max_init_blocks;
    void* r;
    void* z;
    void* p;   // direction buffers: iterations alternate which one holds p (fused path)
    void* p2;
    void* Ap;
};

// Largest row grid (the initial residual): one workgroup per 256/G block rows, G <= 64.
int cg_max_blocks(int64_t nrow) { return (int)std::max<int64_t>(1, ceil_div(nrow, 4)); }

CgWs carve_cg(void* base, size_t bytes, int64_t nrow, int64_t n, size_t scalar_bytes, size_t* needed)
{
    Carve c(base, bytes);
    CgWs w;
    w.max_init_blocks = cg_max_blocks(nrow);
    w.state = c.take<CgState>(1);
    w.init_partials = c.take<double>((size_t)3 * w.max_init_blocks);
    w.spmv_partials = c.take<double>((size_t)kMaxSpmvGrid);
    w.upd_partials = c.take<double>((size_t)2 * kMaxVecGrid);
    w.r = c.take<char>((size_t)n * scalar_bytes);
    w.z = c.take<char>((size_t)n * scalar_bytes);
    w.p = c.take<char>((size_t)n * scalar_bytes);
    w.p2 = c.take<char>((size_t)n * scalar_bytes);
    w.Ap = c.take<char>((size_t)n * scalar_bytes);
    *needed = c.used;
    return w;
}

constexpr int kFinalizeBlock = 1024;

// Fixed-order sum of NV rows of `count` <= MAXP partials (row v at parts + v *
// stride) by one workgroup of kBlock threads, result in every thread.  Every
// workgroup of a kernel that calls it gets bit-identical sums: thread t adds
// parts[t], parts[t + kBlock], ... in order, then the same block reduction.
template <int NV, int MAXP>
__device__ __forceinline__ void sum_parts(const double* __restrict__ parts, size_t stride, int count, double* lds,
                                          double (&out)[NV])
{
    constexpr int kPer = MAXP / kBlock;
    static_assert(MAXP % kBlock == 0, "partial capacity is a multiple of the block");
    double t[NV][kPer];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = (int)threadIdx.x + u * kBlock;
            t[v][u] = k < count ? parts[(size_t)v * stride + k] : 0.0;
        }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            a += t[v][u];
        out[v] = block_sum<kBlock>(a, lds);
    }
}

// Direction buffers.  Unfused, p stays in ws.p.  Fused, iteration k reads
// buffer k & 1 and writes the other: the choice is made on the device, so a
// captured iteration replays correctly any number of times.
template <typename S> __device__ __forceinline__ S* dir_buffer(const CgWs& ws, int which)
{
    return static_cast<S*>(which ? ws.p2 : ws.p);
}

// Vectors up to this size fold the direction pass into the SpMV (measured on
// MI355X: 2D Poisson, 16,129 dofs, 18 -> 15 us per iteration; 823,875-dof
// elasticity and 2M-dof FP64 systems slower by 2-12 us per iteration).
constexpr size_t kFusedDirectionMaxBytes = (size_t)1 << 20;

// z = M r for one block row.
template <typename S, int BR, int PREC>
__device__ __forceinline__ void apply_prec(const S* __restrict__ prec, int row, const S (&r)[BR], S (&z)[BR])
{
    if (PREC == kPrecDiag) {
#pragma unroll
        for (int s = 0; s < BR; ++s)
            z[s] = prec[(size_t)row * BR + s] * r[s];
    } else if (PREC == kPrecBlock) {
        const S* m = prec + (size_t)row * BR * BR;
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            S v = S(0);
#pragma unroll
            for (int c = 0; c < BR; ++c)
                v = fma(m[s * BR + c], r[c], v);
            z[s] = v;
        }
    } else {
#pragma unroll
        for (int s = 0; s < BR; ++s)
            z[s] = r[s];
    }
}

// r = b - A x ; z = M r ; p = z ; partials of (b.b, r.r, r.z).  The grid
// covers the matrix: kCgRpg consecutive block rows per group of G lanes.
template <typename S, int BR, int G, int PREC, int kCgRpg = rows_per_group_for(BR, BR)>
__global__ __launch_bounds__(kBlock) void cg_init_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ columns, const S* __restrict__ values,
    const S* __restrict__ prec, const S* __restrict__ b, const S* __restrict__ x, S* __restrict__ r, S* __restrict__ z,
    int fused, CgWs ws
)
{
    __shared__ double lds[kBlock / kWave];
    S* __restrict__ p = dir_buffer<S>(ws, fused ? (ws.state->iter & 1) : 0);
    const int row0 = (blockIdx.x * (kBlock / G) + threadIdx.x / G) * kCgRpg;
    const int lane = threadIdx.x % G;
    double acc3[3] = {0.0, 0.0, 0.0};
    if (row0 < nrow) {
        const RowAcc<S, BR * kCgRpg> ax =
            group_rows_product<S, BR, BR, G, kCgRpg>(row0, nrow, lane, offsets, nullptr, columns, values, x);
        for (int j = lane; j < kCgRpg && row0 + j < nrow; j += G) {
            const int row = row0 + j;
            S rv[BR], zv[BR];
#pragma unroll
            for (int s = 0; s < BR; ++s) {
                const S bv = b[(size_t)row * BR + s];
                rv[s] = bv - pick(ax, j * BR + s);
                acc3[0] += (double)bv * (double)bv;
            }
            apply_prec<S, BR, PREC>(prec, row, rv, zv);
#pragma unroll
            for (int s = 0; s < BR; ++s) {
                const size_t i = (size_t)row * BR + s;
                r[i] = rv[s];
                if (PREC != kPrecNone)
                    z[i] = zv[s];
                p[i] = zv[s];
                acc3[1] += (double)rv[s] * (double)rv[s];
                acc3[2] += (double)rv[s] * (double)zv[s];
            }
        }
    }
    write_block_partials<kBlock, 3>(acc3, ws.init_partials, lds);
}

// One workgroup: sums the initial residual's partials and (re)initialises
// the state; on the fused path also leaves r.r and r.z as the update's
// partials (first entry, zeros after it), which the first SpMV sums back
// exactly.
__global__ __launch_bounds__(kFinalizeBlock) void cg_init_finalize_kernel(CgWs ws, int nparts, int upd_parts,
                                                                          int has_prec, int first, int fused)
{
    __shared__ double lds[kFinalizeBlock / kWave];
    CgState* state = ws.state;
    double tot[3];
    sum_partials_rows<kFinalizeBlock, 3>(ws.init_partials, (size_t)nparts, nparts, lds, tot);
    const double rz = has_prec ? tot[2] : tot[1];
    if (fused) {
        for (int k = threadIdx.x; k < upd_parts; k += kFinalizeBlock) {
            ws.upd_partials[k] = k == 0 ? tot[1] : 0.0;
            ws.upd_partials[upd_parts + k] = k == 0 ? rz : 0.0;
        }
    }
    if (threadIdx.x != 0)
        return;
    if (first) {
        state->bb = tot[0];
        const double a = fmax(state->rtol * sqrt(tot[0]), state->atol);
        state->atol_sq = a * a;
        state->iter = 0;
    }
    state->rr = tot[1];
    state->rz = rz;
    state->rz_prev = __builtin_inf();
    state->done = !(tot[1] > state->atol_sq);
    state->done_prev = state->done;
    state->alpha = 0.0;
    state->beta = 0.0;
}

// p' = z + beta p, entry i.
template <typename S> struct NextDirection {
    const S* __restrict__ z;
    const S* __restrict__ p;
    S beta;
    __device__ __forceinline__ S operator()(size_t i) const { return fma(beta, p[i], z[i]); }
};

// Ap = A p and p.Ap partials, the grid walking tiles of kBlock / G * kCgRpg
// block rows.  FUSED: the prologue first sums the update's partials into
// r.r and r.z of the current iterate, hence beta and `done` (workgroup 0
// records them), and the product is A p' with p' = z + beta p evaluated
// where it is gathered and written to the other direction buffer.
// Unfused: `done` and p come from the direction pass.
template <typename S, int BR, int G, bool FUSED, int kCgRpg = rows_per_group_for(BR, BR)>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(
    int nrow, const int* __restrict__ offsets, const int* __restrict__ columns, const S* __restrict__ values,
    const S* __restrict__ z, S* __restrict__ Ap, int upd_parts, CgWs ws, double* __restrict__ parts
)
{
    __shared__ double lds[kBlock / kWave];
    CgState* state = ws.state;
    S beta = S(0);
    int par = 0;
    if constexpr (FUSED) {
        double t[2];
        sum_parts<2, kMaxVecGrid>(ws.upd_partials, (size_t)upd_parts, upd_parts, lds, t);
        const bool live = t[0] > state->atol_sq;
        const double bt = live ? t[1] / state->rz_prev : 0.0;
        par = state->iter & 1;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            state->rr = t[0];
            state->rz = t[1];
            state->beta = bt;
            state->done = !live;
            state->iter_copy = state->iter;
        }
        if (!live)
            return;
        beta = (S)bt;
    } else {
        if (state->done)
            return;
    }
    const S* __restrict__ p = dir_buffer<S>(ws, par);
    S* __restrict__ p_next = dir_buffer<S>(ws, par ^ 1);
    const NextDirection<S> dir{z, p, beta};
    constexpr int kTileRows = (kBlock / G) * kCgRpg;
    const int ntiles = (nrow + kTileRows - 1) / kTileRows;
    const int lane = threadIdx.x % G;
    // odd iterations walk the tiles from the last: each product starts with
    // the rows the previous one read last, still in the memory-side cache
    const bool reverse = state->iter & 1;
    double acc[1] = {0.0};
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int tile = reverse ? ntiles - 1 - t : t;
        const int row0 = (tile * (kBlock / G) + (int)threadIdx.x / G) * kCgRpg;
        if (row0 >= nrow)
            continue;
        if constexpr (FUSED) {
            const RowAcc<S, BR * kCgRpg> ap = group_rows_product_f<S, BR, BR, G, kCgRpg, S>(
                row0, nrow, lane, offsets, nullptr, columns, values, dir
            );
            for (int e = lane; e < BR * kCgRpg && row0 + e / BR < nrow; e += G) {
                const size_t i = (size_t)row0 * BR + e;
                const S v = pick(ap, e);
                const S pn = dir(i);
                p_next[i] = pn;
                Ap[i] = v;
                acc[0] += (double)pn * (double)v;
            }
        } else {
            const RowAcc<S, BR * kCgRpg> ap =
                group_rows_product<S, BR, BR, G, kCgRpg>(row0, nrow, lane, offsets, nullptr, columns, values, p);
            for (int e = lane; e < BR * kCgRpg && row0 + e / BR < nrow; e += G) {
                const size_t i = (size_t)row0 * BR + e;
                const S v = pick(ap, e);
                Ap[i] = v;
                acc[0] += (double)p[i] * (double)v;
            }
        }
    }
    write_block_partials<kBlock, 1>(acc, parts, lds);
}

// Wide SpMV grids (one workgroup per tile, more than kMaxSpmvGrid): their
// partials reduced to kWideParts, workgroup b summing the b-th contiguous
// range in a fixed order into spmv_partials[b]; the update sums those.
constexpr int kWideParts = 64;
__global__ __launch_bounds__(kFinalizeBlock) void cg_wide_spmv_finalize_kernel(const double* __restrict__ parts,
                                                                              int nparts, CgWs ws)
{
    __shared__ double lds[kFinalizeBlock / kWave];
    if (ws.state->done)
        return;
    const int chunk = (nparts + kWideParts - 1) / kWideParts;
    const int lo = min((int)blockIdx.x * chunk, nparts);
    const int hi = min(lo + chunk, nparts);
    const double t = sum_partials<kFinalizeBlock>(parts + lo, hi - lo, lds);
    if (threadIdx.x == 0)
        ws.spmv_partials[blockIdx.x] = t;
}

// x += alpha p, r -= alpha Ap, z = M r, partials of (r.r, r.z).  The
// prologue sums the SpMV's partials into p.Ap and forms alpha; workgroup 0
// records alpha and the values the next kernel must read unchanged (r.z of
// this iterate, `done`, the iteration count).  Each thread loads its first
// rows before that sum, so the L2 round trips of the prologue overlap the
// HBM loads.
template <typename S, int BR, int PREC>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(
    int nrow, const S* __restrict__ prec, S* __restrict__ x, S* __restrict__ r, S* __restrict__ z,
    const S* __restrict__ Ap, int fused, int spmv_parts, CgWs ws
)
{
    __shared__ double lds[kBlock / kWave];
    CgState* state = ws.state;
    const int done = state->done;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state->rz_prev = state->rz;
        state->done_prev = done;
        if (!done)
            state->iter = (fused ? state->iter_copy : state->iter) + 1;
    }
    if (done)
        return;
    // the direction this iteration's SpMV used (fused: the buffer it wrote)
    const S* __restrict__ p = dir_buffer<S>(ws, fused ? ((state->iter_copy & 1) ^ 1) : 0);
    const int stride = (int)gridDim.x * kBlock;
    int row = (int)(blockIdx.x * kBlock + threadIdx.x);
    S pv[BR], apv[BR], xv[BR], rv0[BR];
    if (row < nrow) {
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            const size_t i = (size_t)row * BR + s;
            pv[s] = p[i];
            apv[s] = Ap[i];
            xv[s] = x[i];
            rv0[s] = r[i];
        }
    }
    double t[1];
    sum_parts<1, kMaxSpmvGrid>(ws.spmv_partials, 0, spmv_parts, lds, t);
    const double a = state->rr > state->atol_sq ? state->rz / t[0] : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state->pAp = t[0];
        state->alpha = a;
    }
    const S alpha = (S)a;
    double acc[2] = {0.0, 0.0};
    for (bool first = true; row < nrow; row += stride, first = false) {
        if (!first) {
#pragma unroll
            for (int s = 0; s < BR; ++s) {
                const size_t i = (size_t)row * BR + s;
                pv[s] = p[i];
                apv[s] = Ap[i];
                xv[s] = x[i];
                rv0[s] = r[i];
            }
        }
        S rv[BR], zv[BR];
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            const size_t i = (size_t)row * BR + s;
            x[i] = fma(alpha, pv[s], xv[s]);
            rv[s] = fma(-alpha, apv[s], rv0[s]);
            r[i] = rv[s];
        }
        apply_prec<S, BR, PREC>(prec, row, rv, zv);
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            if (PREC != kPrecNone)
                z[(size_t)row * BR + s] = zv[s];
            acc[0] += (double)rv[s] * (double)rv[s];
            acc[1] += (double)rv[s] * (double)zv[s];
        }
    }
    write_block_partials<kBlock, 2>(acc, ws.upd_partials, lds);
}

// p = z + beta p (unfused path).  The prologue sums the update's partials
// into r.r and r.z of the new iterate, forms beta and `done`; workgroup 0
// records them.  z and p are 256-byte aligned workspace slices, moved in
// 16-byte vectors (the first ones loaded before the prologue, the next ones
// while the current are stored); one fma per entry, as a scalar loop.
template <typename S>
__global__ __launch_bounds__(kBlock) void cg_direction_kernel(int64_t n, const S* __restrict__ z, S* __restrict__ p,
                                                              int upd_parts, CgWs ws)
{
    constexpr int kV = 16 / (int)sizeof(S);
    typedef S V __attribute__((ext_vector_type(kV)));
    __shared__ double lds[kBlock / kWave];
    CgState* state = ws.state;
    if (state->done_prev)
        return;
    const V* __restrict__ zv = reinterpret_cast<const V*>(z);
    V* __restrict__ pv = reinterpret_cast<V*>(p);
    const int64_t nv = n / kV;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    V zc = {}, pc = {};
    if (i < nv) {
        zc = zv[i];
        pc = pv[i];
    }
    double t[2];
    sum_parts<2, kMaxVecGrid>(ws.upd_partials, (size_t)upd_parts, upd_parts, lds, t);
    const bool live = t[0] > state->atol_sq;
    const double bt = live ? t[1] / state->rz_prev : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state->rr = t[0];
        state->rz = t[1];
        state->beta = bt;
        state->done = !live;
    }
    const S beta = (S)bt;
    for (; i < nv; i += stride) {
        V zn = {}, pn = {};
        if (i + stride < nv) {
            zn = zv[i + stride];
            pn = pv[i + stride];
        }
        V o;
#pragma unroll
        for (int k = 0; k < kV; ++k)
            o[k] = fma(beta, pc[k], zc[k]);
        pv[i] = o;
        zc = zn;
        pc = pn;
    }
    // the last n % kV entries
    const int64_t j = nv * kV + threadIdx.x;
    if (blockIdx.x == 0 && j < n)
        p[j] = fma(beta, p[j], z[j]);
}

// Fused path at a residual check: r.r, r.z and `done` of the current iterate
// from the update's partials (the next SpMV recomputes the same sums).
__global__ __launch_bounds__(kBlock) void cg_check_kernel(int upd_parts, CgWs ws)
{
    __shared__ double lds[kBlock / kWave];
    CgState* state = ws.state;
    double t[2];
    sum_parts<2, kMaxVecGrid>(ws.upd_partials, (size_t)upd_parts, upd_parts, lds, t);
    if (threadIdx.x == 0) {
        state->rr = t[0];
        state->rz = t[1];
        state->done = !(t[0] > state->atol_sq);
    }
}

__global__ void cg_state_init_kernel(CgWs ws, double rtol, double atol)
{
    CgState* state = ws.state;
    state->rtol = rtol;
    state->atol = atol;
    state->atol_sq = 0.0;
    state->done = 0;
    state->done_prev = 0;
    state->iter = 0;
    state->iter_copy = 0;
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

// Tolerance defaults (linear.py:2010-2032); NaN stands for None.
void get_tolerances(bool is_f64, double tol, double atol, double* rtol_out, double* atol_out)
{
    const double eps = is_f64 ? 1.0e-16 : 1.0e-8;
    const double default_tol = std::pow(eps, 0.75);
    const double min_tol = std::pow(eps, 2.25);
    if (std::isnan(tol) && std::isnan(atol)) {
        tol = atol = default_tol;
    } else if (std::isnan(tol)) {
        tol = atol;
    } else if (std::isnan(atol)) {
        atol = tol;
    }
    *rtol_out = tol;
    *atol_out = std::max(atol, min_tol);
}

// Pinned staging for state read-backs; one per host thread, kept for the
// life of the process (freeing it at thread exit could race runtime teardown).
struct PinnedState {
    CgState* host = nullptr;
};

int read_state(const CgState* dev, CgState* out, hipStream_t stream)
{
    static thread_local PinnedState pinned;
    if (!pinned.host)
        WA_HIP(hipHostMalloc((void**)&pinned.host, sizeof(CgState), hipHostMallocDefault));
    WA_HIP(hipMemcpyAsync(pinned.host, dev, sizeof(CgState), hipMemcpyDeviceToHost, stream));
    WA_HIP(hipStreamSynchronize(stream));
    *out = *pinned.host;
    return kOk;
}

int group_for_square(int64_t nnz_hint, int nrow, int br) { return bsr_mv_lanes(nrow, br, br, nnz_hint); }

typedef int (*CgCallback)(int iteration, double err, double atol, void* user);

struct CgProblem {
    int nrow;
    int br;
    const int* offsets;
    const int* columns;
    const void* values;
    int64_t nnz_hint;
    int prec_kind;
    const void* prec;
    const void* b;
    void* x;
};

template <typename S> struct CgLauncher {
    CgProblem pb;
    CgWs ws;
    hipStream_t stream;
    // Direction pass folded into the SpMV (small systems).
    bool fused = false;

    const S* z_vec() const { return pb.prec_kind != kPrecNone ? (const S*)ws.z : (const S*)ws.r; }

    template <int BR, int G> unsigned rows_grid() const
    {
        return (unsigned)ceil_div(pb.nrow, (kBlock / G) * rows_per_group_for(BR, BR));
    }
    // Very large unfused systems: one workgroup per tile, the partials then
    // reduced to kWideParts by a small kernel (256^3 7-point operator, 65,536 tiles:
    // FP32 inner CG iteration 423 -> 391 us); otherwise at most kMaxSpmvGrid
    // persistent workgroups whose partials the update sums in its prologue
    // (64^3 3x3 elasticity, 17,164 tiles: 66 us per iteration, 71 with the wide grid).
    static constexpr unsigned kWideSpmvTiles = 16u * kMaxSpmvGrid;
    template <int BR, int G> bool wide_spmv() const { return !fused && rows_grid<BR, G>() > kWideSpmvTiles; }
    template <int BR, int G> unsigned spmv_grid() const
    {
        return wide_spmv<BR, G>() ? rows_grid<BR, G>() : std::min<unsigned>(rows_grid<BR, G>(), (unsigned)kMaxSpmvGrid);
    }
    // partials the update reads
    template <int BR, int G> unsigned spmv_parts() const
    {
        return wide_spmv<BR, G>() ? (unsigned)kWideParts : spmv_grid<BR, G>();
    }
    static unsigned vec_grid(int64_t count)
    {
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(count, kBlock), kMaxVecGrid));
    }
    // workgroups (= partials) of the update pass: one thread per block row
    unsigned upd_grid() const { return vec_grid(pb.nrow); }

    template <int BR, int G, int PREC> int init_t(int first)
    {
        const unsigned grid = rows_grid<BR, G>();
        cg_init_kernel<S, BR, G, PREC><<<grid, kBlock, 0, stream>>>(
            pb.nrow, pb.offsets, pb.columns, (const S*)pb.values, (const S*)pb.prec, (const S*)pb.b,
            (const S*)pb.x, (S*)ws.r, (S*)ws.z, fused ? 1 : 0, ws
        );
        WA_LAUNCH_CHECK();
        cg_init_finalize_kernel<<<1, kFinalizeBlock, 0, stream>>>(
            ws, (int)grid, (int)upd_grid(), pb.prec_kind != kPrecNone ? 1 : 0, first, fused ? 1 : 0
        );
        WA_LAUNCH_CHECK();
        return kOk;
    }
    template <int BR, int G> int spmv_t()
    {
        const unsigned grid = spmv_grid<BR, G>();
        if (fused) {
            cg_spmv_kernel<S, BR, G, true><<<grid, kBlock, 0, stream>>>(
                pb.nrow, pb.offsets, pb.columns, (const S*)pb.values, z_vec(), (S*)ws.Ap, (int)upd_grid(), ws,
                ws.spmv_partials
            );
            WA_LAUNCH_CHECK();
            return kOk;
        }
        // the wide grid's partials go to the initial residual's partial buffer
        // (unused between restarts; it holds one per workgroup of any row grid)
        double* parts = wide_spmv<BR, G>() ? ws.init_partials : ws.spmv_partials;
        cg_spmv_kernel<S, BR, G, false><<<grid, kBlock, 0, stream>>>(
            pb.nrow, pb.offsets, pb.columns, (const S*)pb.values, nullptr, (S*)ws.Ap, (int)upd_grid(), ws, parts
        );
        WA_LAUNCH_CHECK();
        if (wide_spmv<BR, G>()) {
            cg_wide_spmv_finalize_kernel<<<kWideParts, kFinalizeBlock, 0, stream>>>(parts, (int)grid, ws);
            WA_LAUNCH_CHECK();
        }
        return kOk;
    }
    template <int BR, int PREC> int update_t()
    {
        unsigned spmv_parts = 0;
#define WA_SPMV_PARTS(BR_, G) spmv_parts = this->template spmv_parts<BR_, G>(); break;
        switch (g()) {
        case 1: WA_SPMV_PARTS(BR, 1)
        case 2: WA_SPMV_PARTS(BR, 2)
        case 4: WA_SPMV_PARTS(BR, 4)
        case 8: WA_SPMV_PARTS(BR, 8)
        case 16: WA_SPMV_PARTS(BR, 16)
        case 32: WA_SPMV_PARTS(BR, 32)
        default: WA_SPMV_PARTS(BR, 64)
        }
#undef WA_SPMV_PARTS
        cg_update_kernel<S, BR, PREC><<<upd_grid(), kBlock, 0, stream>>>(
            pb.nrow, (const S*)pb.prec, (S*)pb.x, (S*)ws.r, (S*)ws.z, (const S*)ws.Ap, fused ? 1 : 0, (int)spmv_parts,
            ws
        );
        WA_LAUNCH_CHECK();
        return kOk;
    }

    int g() const { return group_for_square(pb.nnz_hint, pb.nrow, pb.br); }

#define WA_CG_DISPATCH_G(BR, CALL)                                                                                    \
    switch (g()) {                                                                                                    \
    case 1: return CALL(BR, 1);                                                                                       \
    case 2: return CALL(BR, 2);                                                                                       \
    case 4: return CALL(BR, 4);                                                                                       \
    case 8: return CALL(BR, 8);                                                                                       \
    case 16: return CALL(BR, 16);                                                                                     \
    case 32: return CALL(BR, 32);                                                                                     \
    default: return CALL(BR, 64);                                                                                     \
    }

    int init(int first)
    {
#define WA_INIT_CALL(BR, G) init_prec<BR, G>(first)
        if (pb.br == 1) {
            WA_CG_DISPATCH_G(1, WA_INIT_CALL)
        }
        WA_CG_DISPATCH_G(3, WA_INIT_CALL)
#undef WA_INIT_CALL
    }
    template <int BR, int G> int init_prec(int first)
    {
        switch (pb.prec_kind) {
        case kPrecDiag: return init_t<BR, G, kPrecDiag>(first);
        case kPrecBlock: return init_t<BR, G, kPrecBlock>(first);
        default: return init_t<BR, G, kPrecNone>(first);
        }
    }
    int spmv()
    {
#define WA_SPMV_CALL(BR, G) spmv_t<BR, G>()
        if (pb.br == 1) {
            WA_CG_DISPATCH_G(1, WA_SPMV_CALL)
        }
        WA_CG_DISPATCH_G(3, WA_SPMV_CALL)
#undef WA_SPMV_CALL
    }
    int update()
    {
        if (pb.br == 1) {
            switch (pb.prec_kind) {
            case kPrecDiag: return update_t<1, kPrecDiag>();
            case kPrecBlock: return update_t<1, kPrecBlock>();
            default: return update_t<1, kPrecNone>();
            }
        }
        switch (pb.prec_kind) {
        case kPrecDiag: return update_t<3, kPrecDiag>();
        case kPrecBlock: return update_t<3, kPrecBlock>();
        default: return update_t<3, kPrecNone>();
        }
    }
#undef WA_CG_DISPATCH_G

    int direction()
    {
        const int64_t n = (int64_t)pb.nrow * pb.br;
        cg_direction_kernel<S><<<vec_grid(n), kBlock, 0, stream>>>(n, z_vec(), (S*)ws.p, (int)upd_grid(), ws);
        WA_LAUNCH_CHECK();
        return kOk;
    }

    // Makes the state's r.r and `done` current before the host reads them
    // (the fused path leaves the last update's sums in its partials).
    int check()
    {
        if (!fused)
            return kOk;
        cg_check_kernel<<<1, kBlock, 0, stream>>>((int)upd_grid(), ws);
        WA_LAUNCH_CHECK();
        return kOk;
    }

    // last_in_cycle: a restart follows, which sets p = z itself.
    int iteration(bool last_in_cycle)
    {
        WA_TRY(spmv());
        WA_TRY(update());
        if (!fused && !last_in_cycle)
            WA_TRY(direction());
        return kOk;
    }
};

// Runs the solver's cycles on a private stream and, with graphs on, replays
// each group of cycles between two residual checks as one HIP graph (the
// reference captures its iterations in CUDA graphs, linear.py:2205-2215).  The caller's stream may be the legacy
// default stream, which cannot be captured, hence the private stream; it
// starts after the caller's pending work and the caller's stream waits for
// it at the end.  A failed capture falls back to eager launches.
struct CycleRunner {
    hipStream_t caller;
    hipStream_t ps = nullptr;
    hipEvent_t ev = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_ok = false;

    CycleRunner(hipStream_t s, bool use_graph) : caller(s)
    {
        if (!use_graph)
            return;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
            (void)hipGetLastError();
            return;  // already inside the caller's capture: launch into it
        }
        if (hipStreamCreateWithFlags(&ps, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(ev, caller) != hipSuccess || hipStreamWaitEvent(ps, ev, 0) != hipSuccess) {
            (void)hipGetLastError();
            release();
            return;
        }
        graph_ok = true;
    }
    ~CycleRunner()
    {
        if (ps) {
            // everything enqueued here precedes the caller's later work; the
            // graph is destroyed only once it has run
            if (hipEventRecord(ev, ps) == hipSuccess)
                (void)hipStreamWaitEvent(caller, ev, 0);
            (void)hipStreamSynchronize(ps);
        }
        release();
    }
    void release()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (ps)
            (void)hipStreamDestroy(ps);
        if (ev)
            (void)hipEventDestroy(ev);
        exec = nullptr;
        ps = nullptr;
        ev = nullptr;
    }
    hipStream_t stream() const { return ps ? ps : caller; }

    // `count` cycles of `body`; with graph = true they are one replayable
    // graph (always the same count: the group between two residual checks).
    template <typename F> int run(F&& body, int count, bool graph)
    {
        auto eager = [&]() -> int {
            for (int c = 0; c < count; ++c)
                WA_TRY(body());
            return kOk;
        };
        if (!graph || !graph_ok)
            return eager();
        if (exec) {
            WA_HIP(hipGraphLaunch(exec, ps));
            return kOk;
        }
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(ps, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            graph_ok = false;
            return eager();
        }
        const int st = eager();
        const hipError_t e = hipStreamEndCapture(ps, &g);
        if (st != kOk) {
            if (g)
                (void)hipGraphDestroy(g);
            return st;
        }
        hipError_t ie = e;
        if (e == hipSuccess && g)
            ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (g)
            (void)hipGraphDestroy(g);
        if (e != hipSuccess || !g || ie != hipSuccess) {
            (void)hipGetLastError();
            exec = nullptr;
            graph_ok = false;
            return eager();  // nothing of the capture ran: run the cycles eagerly
        }
        WA_HIP(hipGraphLaunch(exec, ps));
        return kOk;
    }
};

template <typename S>
int cg_solve_impl(
    const CgProblem& pb, double tol, double atol, int maxiter, int check_every, int restart, int use_graph,
    CgCallback cb, void* cb_user, void* workspace, size_t workspace_bytes, int* iters_out, double* rnorm_out,
    double* atol_out, hipStream_t caller_stream
)
{
    WA_REQUIRE(pb.br == 1 || pb.br == 3, "cg_solve: block size must be 1 or 3, got %d", pb.br);
    WA_REQUIRE(pb.prec_kind >= 0 && pb.prec_kind <= 2, "cg_solve: invalid preconditioner kind %d", pb.prec_kind);
    WA_REQUIRE(pb.prec_kind == kPrecNone || pb.prec, "cg_solve: null preconditioner data");
    const int64_t n = (int64_t)pb.nrow * pb.br;
    WA_REQUIRE(pb.nrow > 0, "cg_solve: empty system");
    size_t needed = 0;
    CgWs ws = carve_cg(workspace, workspace_bytes, pb.nrow, n, sizeof(S), &needed);
    WA_REQUIRE(workspace && needed <= workspace_bytes, "cg_solve: workspace too small (%zu < %zu bytes)",
               workspace_bytes, needed);
    // the vector passes move workspace slices in 16-byte vectors
    WA_REQUIRE((uintptr_t)workspace % 16 == 0, "cg_solve: workspace must be 16-byte aligned");
    if (maxiter <= 0)
        maxiter = (int)n;

    // graphs only where launches bound the iteration: measured on MI355X,
    // replaying the cycles between two checks as one graph gave 2D Poisson
    // (16 K dofs) 19 -> 17 us per iteration but the 3x3 elasticity system
    // (824 K dofs) 78.6 -> 79.6 us, and one graph per iteration was slower for both
    CycleRunner runner(caller_stream, use_graph != 0 && (size_t)n * sizeof(S) <= kFusedDirectionMaxBytes);
    const hipStream_t stream = runner.stream();
    CgLauncher<S> L{pb, ws, stream};
    L.fused = (size_t)n * sizeof(S) <= kFusedDirectionMaxBytes;

    int cycle = 1;
    if (restart > 0) {
        cycle = std::max(1, std::min(restart, maxiter));
        if (check_every > 0)
            check_every = std::max(check_every, cycle);
    }

    double rtol_v, atol_v;
    get_tolerances(sizeof(S) == 8, tol, atol, &rtol_v, &atol_v);
    cg_state_init_kernel<<<1, 1, 0, stream>>>(ws, rtol_v, atol_v);
    WA_LAUNCH_CHECK();
    WA_TRY(L.init(1));

    auto cycle_body = [&]() -> int {
        for (int k = 0; k < cycle; ++k)
            WA_TRY(L.iteration(restart > 0 && k == cycle - 1));
        if (restart > 0)
            WA_TRY(L.init(0));
        return kOk;
    };
    // k cycles, then the state made current for the host's read-back
    auto group_of = [&](int k) {
        return [&, k]() -> int {
            for (int c = 0; c < k; ++c)
                WA_TRY(cycle_body());
            return L.check();
        };
    };
    auto run_cycle = [&]() -> int { return runner.run(group_of(1), 1, false); };

    if (check_every <= 0) {
        // Device-side convergence only: enqueue every cycle, no host read-back
        // (groups of kDeviceGroup cycles, one graph replay each).
        constexpr int kDeviceGroup = 16;
        const int total = (int)ceil_div(maxiter, cycle);
        for (int c = 0; c < total; c += kDeviceGroup) {
            const int k = std::min(kDeviceGroup, total - c);
            WA_TRY(runner.run(group_of(k), 1, k == kDeviceGroup));
        }
        if (iters_out || rnorm_out || atol_out) {
            CgState h;
            WA_TRY(read_state(ws.state, &h, stream));
            if (iters_out)
                *iters_out = h.iter;
            if (rnorm_out)
                *rnorm_out = std::sqrt(h.rr);
            if (atol_out)
                *atol_out = std::sqrt(h.atol_sq);
        }
        return kOk;
    }

    CgState h;
    WA_TRY(read_state(ws.state, &h, stream));
    const double atol_out_v = std::sqrt(h.atol_sq);
    int cur_iter = 0;
    if (cb)
        cb(cur_iter, std::sqrt(h.rr), atol_out_v, cb_user);
    if (!(h.rr <= h.atol_sq) && check_every % cycle == 0) {
        // the cycles between two checks run as one group (one graph replay)
        const int per_check = check_every / cycle;
        for (;;) {
            const int left = (int)ceil_div(maxiter - cur_iter, cycle);
            const int k = std::min(per_check, left);
            WA_TRY(runner.run(group_of(k), 1, k == per_check));
            cur_iter += k * cycle;
            if (cur_iter >= maxiter)
                break;
            WA_TRY(read_state(ws.state, &h, stream));
            if (h.rr <= h.atol_sq)
                break;
            if (cb)
                cb(cur_iter, std::sqrt(h.rr), atol_out_v, cb_user);
        }
        WA_TRY(read_state(ws.state, &h, stream));
        if (cb)
            cb(cur_iter, std::sqrt(h.rr), atol_out_v, cb_user);
    } else if (!(h.rr <= h.atol_sq)) {
        for (;;) {
            WA_TRY(run_cycle());
            cur_iter += cycle;
            if (cur_iter >= maxiter)
                break;
            if ((cur_iter % check_every) < cycle) {
                WA_TRY(read_state(ws.state, &h, stream));
                if (h.rr <= h.atol_sq)
                    break;
                if (cb)
                    cb(cur_iter, std::sqrt(h.rr), atol_out_v, cb_user);
            }
        }
        WA_TRY(read_state(ws.state, &h, stream));
        if (cb)
            cb(cur_iter, std::sqrt(h.rr), atol_out_v, cb_user);
    }
    if (iters_out)
        *iters_out = cur_iter;
    if (rnorm_out)
        *rnorm_out = std::sqrt(h.rr);
    if (atol_out)
        *atol_out = atol_out_v;
    return kOk;
}

// ---------------------------------------------------------------------------
// Generic primitives (operators supplied by the caller)
// ---------------------------------------------------------------------------

template <typename S>
__global__ __launch_bounds__(kBlock) void dot_partials_kernel(
    int64_t n, const S* __restrict__ a, const S* __restrict__ b, double* partials
)
{
    __shared__ double lds[kBlock / kWave];
    double acc[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        acc[0] += (double)a[i] * (double)b[i];
    write_block_partials<kBlock, 1>(acc, partials, lds);
}

__global__ __launch_bounds__(kFinalizeBlock) void dot_finalize_kernel(const double* partials, int nparts, double* out)
{
    __shared__ double lds[kFinalizeBlock / kWave];
    const double t = sum_partials<kFinalizeBlock>(partials, nparts, lds);
    if (threadIdx.x == 0)
        *out = t;
}

// x += alpha p ; r -= alpha Ap with alpha = rr > atol_sq ? rz / pAp : 0
template <typename S>
__global__ __launch_bounds__(kBlock) void step_xr_kernel(
    int64_t n, S* __restrict__ x, S* __restrict__ r, const S* __restrict__ p, const S* __restrict__ Ap,
    const double* rr, const double* atol_sq, const double* rz, const double* pAp
)
{
    const double a = *rr > *atol_sq ? *rz / *pAp : 0.0;
    const S alpha = (S)a;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        x[i] = fma(alpha, p[i], x[i]);
        r[i] = fma(-alpha, Ap[i], r[i]);
    }
}

// p = z + beta p with beta = rr > atol_sq ? rz_new / rz_old : 0
template <typename S>
__global__ __launch_bounds__(kBlock) void step_p_kernel(
    int64_t n, const S* __restrict__ z, S* __restrict__ p, const double* rr, const double* atol_sq,
    const double* rz_new, const double* rz_old
)
{
    const double bt = *rr > *atol_sq ? *rz_new / *rz_old : 0.0;
    const S beta = (S)bt;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        p[i] = fma(beta, p[i], z[i]);
}

// atol_sq = max(rtol * sqrt(bb), atol)^2 (linear.py:2051-2060)
__global__ void tolerance_kernel(const double* bb, double rtol, double atol, double* atol_sq)
{
    const double a = fmax(rtol * sqrt(*bb), atol);
    *atol_sq = a * a;
}

// ---------------------------------------------------------------------------
// Batched solves (linear.py:33-131, 2010-2373 with batch_offsets): scalar dofs
// [off[b], off[b + 1]) form subproblem b < nb; dofs below off[0] or at or
// beyond off[nb] are inactive.  Dots are per subproblem; alpha / beta and
// the convergence test are per subproblem too.
// ---------------------------------------------------------------------------

// Dofs per workgroup of a segmented dot.  Chunks never cross a subproblem
// boundary, so each subproblem's partials are contiguous and summed in a
// fixed order: the results do not depend on timing.
constexpr int kSegChunk = 4 * kBlock;

__global__ __launch_bounds__(kBlock) void seg_chunk_counts_kernel(int nb, const int* __restrict__ off,
                                                                 int* __restrict__ counts)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b < nb) {
        const int len = off[b + 1] - off[b];
        counts[b] = len > 0 ? (len + kSegChunk - 1) / kSegChunk : 0;
    }
}

// Subproblem of chunk w: the last b with chunk_start[b] <= w.
__device__ __forceinline__ int chunk_batch(int w, int nb, const int* chunk_start)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (chunk_start[mid] <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <typename S>
__global__ __launch_bounds__(kBlock) void seg_dot_partials_kernel(int nb, const int* __restrict__ off,
                                                                 const int* __restrict__ chunk_start,
                                                                 const S* __restrict__ a, const S* __restrict__ c,
                                                                 double* __restrict__ partials)
{
    __shared__ double lds[kBlock / kWave];
    const int w = blockIdx.x;
    if (w >= chunk_start[nb])
        return;  // workgroup-uniform
    const int bt = chunk_batch(w, nb, chunk_start);
    const int64_t beg = off[bt] + (int64_t)(w - chunk_start[bt]) * kSegChunk;
    const int64_t end = min(beg + (int64_t)kSegChunk, (int64_t)off[bt + 1]);
    double acc = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
        acc += (double)a[i] * (double)c[i];
    const double sum = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0)
        partials[w] = sum;
}

// out[b] = sum of subproblem b's partials, one workgroup per subproblem.
__global__ __launch_bounds__(kBlock) void seg_dot_finalize_kernel(const int* __restrict__ chunk_start,
                                                                 const double* __restrict__ partials,
                                                                 double* __restrict__ out)
{
    __shared__ double lds[kBlock / kWave];
    const int bt = blockIdx.x;
    const int c0 = chunk_start[bt], c1 = chunk_start[bt + 1];
    double acc = 0.0;
    for (int k = c0 + (int)threadIdx.x; k < c1; k += kBlock)
        acc += partials[k];
    const double sum = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0)
        out[bt] = sum;
}

// Subproblem of entry e (dpe scalars; the entry's first dof decides, as the
// reference's _find_entry_batch does), -1 when inactive.
__device__ __forceinline__ int entry_batch(int64_t e, int dpe, int nb, const int* off)
{
    const int64_t dof = e * dpe;
    if (dof < off[0] || dof >= off[nb])
        return -1;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (off[mid] <= dof)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <typename S>
__global__ __launch_bounds__(kBlock) void step_xr_batched_kernel(
    int64_t nent, int dpe, int nb, const int* __restrict__ off, S* __restrict__ x, S* __restrict__ r,
    const S* __restrict__ p, const S* __restrict__ Ap, const double* rr, const double* atol_sq, const double* rz,
    const double* pAp
)
{
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nent; e += (int64_t)gridDim.x * kBlock) {
        const int bt = entry_batch(e, dpe, nb, off);
        if (bt < 0)
            continue;
        const S alpha = (S)(rr[bt] > atol_sq[bt] ? rz[bt] / pAp[bt] : 0.0);
        for (int k = 0; k < dpe; ++k) {
            const int64_t i = e * dpe + k;
            x[i] = fma(alpha, p[i], x[i]);
            r[i] = fma(-alpha, Ap[i], r[i]);
        }
    }
}

template <typename S>
__global__ __launch_bounds__(kBlock) void step_p_batched_kernel(
    int64_t nent, int dpe, int nb, const int* __restrict__ off, const S* __restrict__ z, S* __restrict__ p,
    const double* rr, const double* atol_sq, const double* rz_new, const double* rz_old
)
{
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nent; e += (int64_t)gridDim.x * kBlock) {
        const int bt = entry_batch(e, dpe, nb, off);
        if (bt < 0)
            continue;
        const S beta = (S)(rr[bt] > atol_sq[bt] ? rz_new[bt] / rz_old[bt] : 0.0);
        for (int k = 0; k < dpe; ++k) {
            const int64_t i = e * dpe + k;
            p[i] = fma(beta, p[i], z[i]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void tolerance_batched_kernel(int nb, const double* bb, double rtol,
                                                                  double atol, double* atol_sq)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b < nb) {
        const double a = fmax(rtol * sqrt(bb[b]), atol);
        atol_sq[b] = a * a;
    }
}

// z = alpha * (d o x) + beta * y (linear.py:2271-2287)
template <typename S>
__global__ __launch_bounds__(kBlock) void diag_mv_kernel(
    int64_t n, const S* __restrict__ d, const S* __restrict__ x, const S* __restrict__ y, S* __restrict__ z, S alpha,
    S beta
)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        S s = S(0);
        if (alpha != S(0))
            s += alpha * (d[i] * x[i]);
        if (beta != S(0))
            s += beta * y[i];
        z[i] = s;
    }
}

// z = alpha * A x + beta * y for a dense row-major m x k matrix, one wave per
// row (linear.py:2246-2268).
template <typename S>
__global__ __launch_bounds__(kBlock) void dense_mv_kernel(
    int m, int k, const S* __restrict__ A, const S* __restrict__ x, const S* __restrict__ y, S* __restrict__ z,
    S alpha, S beta
)
{
    const int row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= m)
        return;
    S s = S(0);
    if (alpha != S(0))
        for (int c = lane; c < k; c += kWave)
            s += A[(size_t)row * k + c] * x[c];
    s = wave_sum(s) * alpha;
    if (lane == 0) {
        if (beta != S(0))
            s += beta * y[row];
        z[row] = s;
    }
}

int vec_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, ceil_div(n, kBlock))); }

}  // namespace
}  // namespace wa

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// n scalar unknowns; the reduction scratch is sized for n block rows, an
// upper bound for every block size.
WA_API size_t warp_amd_cg_workspace_size(int64_t n, int scalar_size)
{
    size_t needed = 0;
    wa::carve_cg(nullptr, 0, n, n, (size_t)scalar_size, &needed);
    return needed + 256;
}

// Byte offsets of the device-resident scalars inside the CG workspace, so
// callers can read them without a host copy of the layout:
// 0 iterations (int32), 1 r.r (double), 2 atol^2 (double).
WA_API size_t warp_amd_cg_state_offset(int which)
{
    switch (which) {
    case 0: return offsetof(wa::CgState, iter);
    case 1: return offsetof(wa::CgState, rr);
    case 2: return offsetof(wa::CgState, atol_sq);
    default: return 0;
    }
}

#define WA_DEFINE_CG(SUFFIX, S)                                                                                       \
    WA_API int warp_amd_cg_solve_##SUFFIX(                                                                            \
        int nrow, int block_size, const int* offsets, const int* columns, const S* values, int64_t nnz_hint,          \
        int prec_kind, const S* prec, const S* b, S* x, double tol, double atol, int maxiter, int check_every,        \
        int restart, int use_graph, wa::CgCallback callback, void* callback_user, void* workspace,                    \
        size_t workspace_bytes,                                                                                       \
        int* iterations, double* residual_norm, double* abs_tolerance, void* stream                                   \
    )                                                                                                                 \
    {                                                                                                                 \
        wa::CgProblem pb{nrow, block_size, offsets, columns, values, nnz_hint, prec_kind, prec, b, x};                \
        return wa::cg_solve_impl<S>(                                                                                  \
            pb, tol, atol, maxiter, check_every, restart, use_graph, callback, callback_user, workspace,              \
            workspace_bytes, iterations, residual_norm, abs_tolerance, wa::as_stream(stream)                          \
        );                                                                                                            \
    }                                                                                                                 \
    WA_API int warp_amd_dot_##SUFFIX(                                                                                 \
        int64_t n, const S* a, const S* b, double* out, void* workspace, size_t workspace_bytes, void* stream         \
    )                                                                                                                 \
    {                                                                                                                 \
        const int grid = wa::vec_grid(n);                                                                             \
        WA_REQUIRE(workspace && workspace_bytes >= sizeof(double) * grid,                                            \
                   "dot: workspace too small (need %zu bytes)", sizeof(double) * grid);                               \
        double* partials = (double*)workspace;                                                                        \
        wa::dot_partials_kernel<S><<<grid, wa::kBlock, 0, wa::as_stream(stream)>>>(n, a, b, partials);               \
        WA_LAUNCH_CHECK();                                                                                            \
        wa::dot_finalize_kernel<<<1, wa::kFinalizeBlock, 0, wa::as_stream(stream)>>>(partials, grid, out);           \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_cg_step_xr_##SUFFIX(                                                                          \
        int64_t n, S* x, S* r, const S* p, const S* Ap, const double* rr, const double* atol_sq, const double* rz,    \
        const double* pAp, void* stream                                                                               \
    )                                                                                                                 \
    {                                                                                                                 \
        wa::step_xr_kernel<S><<<wa::vec_grid(n), wa::kBlock, 0, wa::as_stream(stream)>>>(                             \
            n, x, r, p, Ap, rr, atol_sq, rz, pAp                                                                      \
        );                                                                                                            \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_cg_step_p_##SUFFIX(                                                                           \
        int64_t n, const S* z, S* p, const double* rr, const double* atol_sq, const double* rz_new,                   \
        const double* rz_old, void* stream                                                                            \
    )                                                                                                                 \
    {                                                                                                                 \
        wa::step_p_kernel<S><<<wa::vec_grid(n), wa::kBlock, 0, wa::as_stream(stream)>>>(                              \
            n, z, p, rr, atol_sq, rz_new, rz_old                                                                      \
        );                                                                                                            \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_seg_dot_##SUFFIX(                                                                             \
        int64_t n, int nb, const int* batch_offsets, const int* chunk_start, int max_chunks, const S* a, const S* b,  \
        double* partials, double* out, void* stream                                                                   \
    )                                                                                                                 \
    {                                                                                                                 \
        WA_REQUIRE(nb > 0 && max_chunks >= 0 && batch_offsets && chunk_start && partials && out,                      \
                   "seg_dot: invalid arguments");                                                                     \
        hipStream_t s = wa::as_stream(stream);                                                                        \
        if (max_chunks > 0) {                                                                                         \
            wa::seg_dot_partials_kernel<S><<<(unsigned)max_chunks, wa::kBlock, 0, s>>>(nb, batch_offsets,            \
                                                                                      chunk_start, a, b, partials);   \
            WA_LAUNCH_CHECK();                                                                                        \
        }                                                                                                             \
        wa::seg_dot_finalize_kernel<<<(unsigned)nb, wa::kBlock, 0, s>>>(chunk_start, partials, out);                  \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_cg_step_xr_batched_##SUFFIX(                                                                  \
        int64_t nent, int dpe, int nb, const int* batch_offsets, S* x, S* r, const S* p, const S* Ap,                 \
        const double* rr, const double* atol_sq, const double* rz, const double* pAp, void* stream                    \
    )                                                                                                                 \
    {                                                                                                                 \
        WA_REQUIRE(nb > 0 && dpe > 0 && batch_offsets, "cg_step_xr_batched: invalid arguments");                     \
        if (nent <= 0)                                                                                                \
            return wa::kOk;                                                                                           \
        wa::step_xr_batched_kernel<S><<<wa::vec_grid(nent), wa::kBlock, 0, wa::as_stream(stream)>>>(                 \
            nent, dpe, nb, batch_offsets, x, r, p, Ap, rr, atol_sq, rz, pAp                                          \
        );                                                                                                            \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_cg_step_p_batched_##SUFFIX(                                                                   \
        int64_t nent, int dpe, int nb, const int* batch_offsets, const S* z, S* p, const double* rr,                  \
        const double* atol_sq, const double* rz_new, const double* rz_old, void* stream                               \
    )                                                                                                                 \
    {                                                                                                                 \
        WA_REQUIRE(nb > 0 && dpe > 0 && batch_offsets, "cg_step_p_batched: invalid arguments");                      \
        if (nent <= 0)                                                                                                \
            return wa::kOk;                                                                                           \
        wa::step_p_batched_kernel<S><<<wa::vec_grid(nent), wa::kBlock, 0, wa::as_stream(stream)>>>(                  \
            nent, dpe, nb, batch_offsets, z, p, rr, atol_sq, rz_new, rz_old                                           \
        );                                                                                                            \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_diag_mv_##SUFFIX(                                                                             \
        int64_t n, const S* d, const S* x, const S* y, S* z, S alpha, S beta, void* stream                            \
    )                                                                                                                 \
    {                                                                                                                 \
        if (n <= 0)                                                                                                   \
            return wa::kOk;                                                                                           \
        wa::diag_mv_kernel<S><<<wa::vec_grid(n), wa::kBlock, 0, wa::as_stream(stream)>>>(n, d, x, y, z, alpha, beta); \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }                                                                                                                 \
    WA_API int warp_amd_dense_mv_##SUFFIX(                                                                            \
        int m, int k, const S* A, const S* x, const S* y, S* z, S alpha, S beta, void* stream                         \
    )                                                                                                                 \
    {                                                                                                                 \
        if (m <= 0)                                                                                                   \
            return wa::kOk;                                                                                           \
        const int rows_per_block = wa::kBlock / wa::kWave;                                                            \
        wa::dense_mv_kernel<S><<<(unsigned)wa::ceil_div(m, rows_per_block), wa::kBlock, 0, wa::as_stream(stream)>>>(  \
            m, k, A, x, y, z, alpha, beta                                                                             \
        );                                                                                                            \
        WA_LAUNCH_CHECK();                                                                                            \
        return wa::kOk;                                                                                               \
    }

WA_DEFINE_CG(f32, float)
WA_DEFINE_CG(f64, double)

WA_API size_t warp_amd_dot_workspace_size(int64_t n) { return sizeof(double) * wa::vec_grid(n); }

// Chunks a segmented dot over n active dofs in nb subproblems can use.
WA_API int warp_amd_seg_dot_max_chunks(int64_t n, int nb)
{
    return (int)(wa::ceil_div(n, wa::kSegChunk) + nb);
}

WA_API size_t warp_amd_seg_dot_plan_workspace_size(int nb)
{
    return wa::align_up(sizeof(int) * (size_t)(nb + 1), 256) + wa::scan_workspace_bytes((int64_t)nb + 1);
}

// chunk_start[0..nb] = exclusive scan of the chunks per subproblem
// (chunk_start[nb] = total); computed once per batch layout.
WA_API int warp_amd_seg_dot_plan(int nb, const int* batch_offsets, int* chunk_start, void* workspace,
                                 size_t workspace_bytes, void* stream)
{
    WA_REQUIRE(nb > 0 && batch_offsets && chunk_start, "seg_dot_plan: invalid arguments");
    WA_REQUIRE(workspace && workspace_bytes >= warp_amd_seg_dot_plan_workspace_size(nb),
               "seg_dot_plan: workspace too small");
    hipStream_t s = wa::as_stream(stream);
    int* counts = static_cast<int*>(workspace);
    void* scan_ws = static_cast<char*>(workspace) + wa::align_up(sizeof(int) * (size_t)(nb + 1), 256);
    wa::seg_chunk_counts_kernel<<<(unsigned)wa::ceil_div(nb, wa::kBlock), wa::kBlock, 0, s>>>(nb, batch_offsets,
                                                                                            counts);
    WA_LAUNCH_CHECK();
    return wa::exclusive_scan_i32(counts, chunk_start, nb, scan_ws, s);
}

WA_API int warp_amd_cg_tolerance_batched(int nb, const double* bb, double tol, double atol, int is_f64,
                                         double* atol_sq, void* stream)
{
    WA_REQUIRE(nb > 0 && bb && atol_sq, "cg_tolerance_batched: invalid arguments");
    double rtol_v, atol_v;
    wa::get_tolerances(is_f64 != 0, tol, atol, &rtol_v, &atol_v);
    wa::tolerance_batched_kernel<<<(unsigned)wa::ceil_div(nb, wa::kBlock), wa::kBlock, 0, wa::as_stream(stream)>>>(
        nb, bb, rtol_v, atol_v, atol_sq);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}

WA_API int warp_amd_cg_tolerance(const double* bb, double tol, double atol, int is_f64, double* atol_sq, void* stream)
{
    double rtol_v, atol_v;
    wa::get_tolerances(is_f64 != 0, tol, atol, &rtol_v, &atol_v);
    wa::tolerance_kernel<<<1, 1, 0, wa::as_stream(stream)>>>(bb, rtol_v, atol_v, atol_sq);
    WA_LAUNCH_CHECK();
    return wa::kOk;
}
