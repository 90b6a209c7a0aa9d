// edm_blas.hip — complex-FP64 vector kernels for the Krylov solver.
//
// These stand in for KrylovKit's BLAS-1 on device arrays (dot, norm, axpy!, scale!,
// orthogonalize) and the chunked restart GEMM of MomentumED/ext/MomentumEDCUDAExt.jl:210-244.
// All are HBM-bound streaming kernels over vectors of D complex doubles laid out
// contiguously (16 B per element). Reductions are two-stage and fixed-order (per-block
// partials, then one block per output summing the partials in block order), so results are
// bitwise reproducible run to run, independent of what else runs on the device.
#include "edm_internal.h"

#include <algorithm>

namespace {

constexpr int kBlk = 256;
constexpr int kRpt = 4;                    // rows per thread in reductions
constexpr int kTile = kBlk * kRpt;         // rows per reduction block
constexpr int kMaxVec = 256;               // vectors per multi-dot / multi-axpy call

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The complex products of the Gram-Schmidt kernels with their rounding fixed in the source
// (one product of each component rounded, the other fused into an explicit fma, no
// contraction left to the compiler, which may fuse either product depending on the code
// around it), so the single-problem and the batched kernels get the same bits for the same
// operands.
// acc -= a * c
__device__ __forceinline__ void cmul_sub(double2 &acc, double2 a, double2 c)
{
#pragma clang fp contract(off)
    acc.x -= fma(a.x, c.x, -(a.y * c.y));
    acc.y -= fma(a.y, c.x, a.x * c.y);
}

// (ar, ai) += conj(a) * w
__device__ __forceinline__ void cdot_acc(double &ar, double &ai, double2 a, double2 w)
{
#pragma clang fp contract(off)
    ar += fma(a.y, w.y, a.x * w.x);
    ai += fma(a.x, w.y, -(a.y * w.x));
}

// block-wide sum of (a, b); result valid in thread 0
__device__ __forceinline__ double2 block_sum2(double a, double b, double2 *sh)
{
    a = wave_sum(a);
    b = wave_sum(b);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = make_double2(a, b);
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { r.x += sh[w].x; r.y += sh[w].y; }
    return r;
}

// The multi-dot of one reduction tile (kTile rows from `base`, D rows in all) for vectors
// i0 <= i < i1: out[i * stride] = sum over the tile's rows of conj(V_i[r]) * w[r]. Per vector
// the sum is the fixed-order block sum of block_sum2 (each thread's kRpt rows in order, the
// wave butterfly, then the waves in order), whichever block computes it; the vectors go
// kGroup at a time, so a group's loads are in flight together and one pair of barriers
// serves the group.
constexpr int kGroup = 4;          // vectors per reduction round
constexpr int kMinDotBlocks = 2048;   // blocks a multi-dot grid should give the device

// Vectors per block of a multi-dot over nb tiles: all of them when the tiles alone fill the
// device, else split (in whole groups) over a second grid dimension.
inline int dot_span(int64_t nb, int nvec)
{
    const int groups = (nvec + kGroup - 1) / kGroup;
    int64_t ny = (kMinDotBlocks + nb - 1) / std::max<int64_t>(nb, 1);
    ny = std::max<int64_t>(1, std::min<int64_t>(ny, groups));
    return kGroup * (int)((groups + ny - 1) / ny);
}

__device__ __forceinline__ void multidot_tile(const double2 *const *__restrict__ vptrs, int i0,
                                              int i1, const double2 *__restrict__ w, int64_t D,
                                              int64_t base, double2 *__restrict__ out,
                                              int64_t stride)
{
    __shared__ double2 sh[kGroup][kBlk / 64];
    double2 wv[kRpt];
    bool in[kRpt];
#pragma unroll
    for (int k = 0; k < kRpt; ++k) {
        int64_t r = base + (int64_t)k * kBlk;
        in[k] = r < D;
        wv[k] = in[k] ? w[r] : make_double2(0.0, 0.0);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int g0 = i0; g0 < i1; g0 += kGroup) {
        double2 a[kGroup][kRpt];
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
            const double2 *v = vptrs[g0 + g < i1 ? g0 + g : g0];
#pragma unroll
            for (int k = 0; k < kRpt; ++k)
                a[g][k] = in[k] ? v[base + (int64_t)k * kBlk] : make_double2(0.0, 0.0);
        }
        double ar[kGroup], ai[kGroup];
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
            ar[g] = 0.0;
            ai[g] = 0.0;
#pragma unroll
            for (int k = 0; k < kRpt; ++k) {
                if (in[k]) cdot_acc(ar[g], ai[g], a[g][k], wv[k]);
            }
            ar[g] = wave_sum(ar[g]);
            ai[g] = wave_sum(ai[g]);
        }
        __syncthreads();   // the previous round's sums have been read
        if (lane == 0) {
#pragma unroll
            for (int g = 0; g < kGroup; ++g) sh[g][wid] = make_double2(ar[g], ai[g]);
        }
        __syncthreads();
        if (threadIdx.x < kGroup && g0 + (int)threadIdx.x < i1) {
            double2 r = make_double2(0.0, 0.0);
            for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
                r.x += sh[threadIdx.x][q].x;
                r.y += sh[threadIdx.x][q].y;
            }
            out[(int64_t)(g0 + threadIdx.x) * stride] = r;
        }
    }
}

// partial[i * nb + blk] = sum over the block's rows of conj(V_i[r]) * w[r]; block (blk, y)
// takes the vectors [y * span, (y + 1) * span)
__global__ void __launch_bounds__(kBlk)
k_multidot_partial(const double2 *const *__restrict__ vptrs, int nvec, int span,
                   const double2 *__restrict__ w, int64_t D, double2 *__restrict__ partial)
{
    const int i0 = (int)blockIdx.y * span, i1 = min(nvec, i0 + span);
    multidot_tile(vptrs, i0, i1, w, D, (int64_t)blockIdx.x * kTile + threadIdx.x,
                  partial + blockIdx.x, gridDim.x);
}

// out[i] = sum_b partial[i * nb + b], one block per output, fixed order
__global__ void __launch_bounds__(kBlk)
k_reduce_partials(const double2 *__restrict__ partial, int64_t nb, double2 *__restrict__ out)
{
    __shared__ double2 sh[kBlk / 64];
    const double2 *p = partial + (int64_t)blockIdx.x * nb;
    double ar = 0.0, ai = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kBlk) { ar += p[b].x; ai += p[b].y; }
    double2 s = block_sum2(ar, ai, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// w[r] -= sum_i V_i[r] h_i
__global__ void __launch_bounds__(kBlk)
k_multiaxpy(const double2 *const *__restrict__ vptrs, int nvec, const double2 *__restrict__ h,
            double2 *__restrict__ w, int64_t D)
{
    __shared__ double2 hs[kMaxVec];
    for (int i = threadIdx.x; i < nvec; i += kBlk) hs[i] = h[i];
    __syncthreads();
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    double2 acc = w[r];
    for (int i = 0; i < nvec; ++i) cmul_sub(acc, vptrs[i][r], hs[i]);
    w[r] = acc;
}

__global__ void __launch_bounds__(kBlk)
k_norm2_partial(const double2 *__restrict__ w, int64_t D, double2 *__restrict__ partial)
{
    __shared__ double2 sh[kBlk / 64];
    int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kRpt; ++k) {
        int64_t r = base + (int64_t)k * kBlk;
        if (r < D) { double2 a = w[r]; s += a.x * a.x + a.y * a.y; }
    }
    double2 t = block_sum2(s, 0.0, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_sqrt_first(const double2 *__restrict__ in, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = sqrt(in[0].x);
}

__global__ void __launch_bounds__(kBlk)
k_scale_inv(double2 *__restrict__ v, const double2 *__restrict__ w, const double *__restrict__ beta,
            int64_t D)
{
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    double inv = 1.0 / beta[0];
    double2 a = w[r];
    v[r] = make_double2(a.x * inv, a.y * inv);
}

// out_c[r] = sum_i V_i[r] * Y[i + ldy*c], c in [c0, c0+kc)
constexpr int kRotChunk = 8;
__global__ void __launch_bounds__(kBlk)
k_rotate(const double2 *const *__restrict__ vptrs, int m, const double *__restrict__ Y, int ldy,
         int c0, int kc, double2 *const *__restrict__ outptrs, int64_t D)
{
    extern __shared__ double ys[];   // m * kRotChunk
    for (int t = threadIdx.x; t < m * kRotChunk; t += kBlk) {
        int i = t / kRotChunk, c = t % kRotChunk;
        ys[t] = c < kc ? Y[i + (int64_t)ldy * (c0 + c)] : 0.0;
    }
    __syncthreads();
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    double ar[kRotChunk], ai[kRotChunk];
#pragma unroll
    for (int c = 0; c < kRotChunk; ++c) { ar[c] = 0.0; ai[c] = 0.0; }
    for (int i = 0; i < m; ++i) {
        double2 v = vptrs[i][r];
#pragma unroll
        for (int c = 0; c < kRotChunk; ++c) {
            double y = ys[i * kRotChunk + c];
            ar[c] += v.x * y;
            ai[c] += v.y * y;
        }
    }
    for (int c = 0; c < kc; ++c) outptrs[c0 + c][r] = make_double2(ar[c], ai[c]);
}

// the same with complex Y (block Lanczos: the projected matrix is complex Hermitian)
__global__ void __launch_bounds__(kBlk)
k_rotate_c(const double2 *const *__restrict__ vptrs, int m, const double2 *__restrict__ Y,
           int ldy, int c0, int kc, double2 *const *__restrict__ outptrs, int64_t D)
{
    extern __shared__ double2 yc[];   // m * kRotChunk
    for (int t = threadIdx.x; t < m * kRotChunk; t += kBlk) {
        int i = t / kRotChunk, c = t % kRotChunk;
        yc[t] = c < kc ? Y[i + (int64_t)ldy * (c0 + c)] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    double ar[kRotChunk], ai[kRotChunk];
#pragma unroll
    for (int c = 0; c < kRotChunk; ++c) { ar[c] = 0.0; ai[c] = 0.0; }
    for (int i = 0; i < m; ++i) {
        double2 v = vptrs[i][r];
#pragma unroll
        for (int c = 0; c < kRotChunk; ++c) {
            double2 y = yc[i * kRotChunk + c];
            ar[c] += v.x * y.x - v.y * y.y;
            ai[c] += v.x * y.y + v.y * y.x;
        }
    }
    for (int c = 0; c < kc; ++c) outptrs[c0 + c][r] = make_double2(ar[c], ai[c]);
}

// SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) as a counter-based generator:
// value(seed, n) = mix(seed + (n + 1) * golden).
__device__ __forceinline__ uint64_t splitmix(uint64_t seed, uint64_t n)
{
    uint64_t z = seed + (n + 1ULL) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }

__global__ void __launch_bounds__(kBlk)
k_fill_random(double2 *__restrict__ v, int64_t D, uint64_t seed, int mode)
{
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    double a = u01(splitmix(seed, 2 * (uint64_t)r));
    double b = u01(splitmix(seed, 2 * (uint64_t)r + 1));
    if (mode == 0) {
        v[r] = make_double2(a, b);
    } else {   // Box-Muller: two independent N(0,1) values
        double rad = sqrt(-2.0 * log(1.0 - a));
        double s, c;
        sincospi(2.0 * b, &s, &c);
        v[r] = make_double2(rad * c, rad * s);
    }
}

// ---------------------------------------------------------------- batched forms
// The same operations for many problems in one launch. A tile (kTile rows) or a row block
// (kBlk rows) belongs to exactly one problem, found through a table; per-problem partial sums
// live in the problem's own region and are reduced in tile order, so every problem's result
// is bitwise what the single-problem kernels above produce, whatever else shares the launch.
__global__ void __launch_bounds__(kBlk)
k_bmultidot_partial(const BatchVecDesc *__restrict__ descs, const int *__restrict__ tile2prob,
                    int span)
{
    const BatchVecDesc &P = descs[tile2prob[blockIdx.x]];
    int64_t lt = (int64_t)blockIdx.x - P.tile0;
    const int i0 = (int)blockIdx.y * span, i1 = min(P.nvec, i0 + span);
    if (i0 >= i1) return;   // uniform over the block
    multidot_tile(P.V, i0, i1, P.w, P.D, lt * kTile + threadIdx.x, P.partial + lt, P.ntiles);
}

// out[i] = sum of partials, block (problem, i); which = 0 -> h, 1 -> h2
__global__ void __launch_bounds__(kBlk)
k_breduce(const BatchVecDesc *__restrict__ descs, int which)
{
    __shared__ double2 sh[kBlk / 64];
    const BatchVecDesc &P = descs[blockIdx.x];
    int i = blockIdx.y;
    if (i >= P.nvec) return;
    const double2 *p = P.partial + (int64_t)i * P.ntiles;
    double ar = 0.0, ai = 0.0;
    for (int64_t b = threadIdx.x; b < P.ntiles; b += kBlk) { ar += p[b].x; ai += p[b].y; }
    double2 s = block_sum2(ar, ai, sh);
    if (threadIdx.x == 0) (which ? P.h2 : P.h)[i] = s;
}

// rev: row blocks taken in reverse dispatch order (block b does the work of block nblk-1-b)
__global__ void __launch_bounds__(kBlk)
k_baxpy(const BatchVecDesc *__restrict__ descs, const int *__restrict__ blk2prob, int which,
        int rev)
{
    __shared__ double2 hs[kMaxVec];
    const int64_t bx = rev ? (int64_t)gridDim.x - 1 - blockIdx.x : (int64_t)blockIdx.x;
    const BatchVecDesc &P = descs[blk2prob[bx]];
    const double2 *c = which ? P.h2 : P.h;
    for (int i = threadIdx.x; i < P.nvec; i += kBlk) hs[i] = c[i];
    __syncthreads();
    int64_t r = (bx - P.blk0) * kBlk + threadIdx.x;
    if (r >= P.D) return;
    double2 acc = P.w[r];
    for (int i = 0; i < P.nvec; ++i) cmul_sub(acc, P.V[i][r], hs[i]);
    P.w[r] = acc;
}

__global__ void __launch_bounds__(kBlk)
k_bnorm_partial(const BatchVecDesc *__restrict__ descs, const int *__restrict__ tile2prob)
{
    __shared__ double2 sh[kBlk / 64];
    const BatchVecDesc &P = descs[tile2prob[blockIdx.x]];
    int64_t lt = (int64_t)blockIdx.x - P.tile0;
    int64_t base = lt * kTile + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kRpt; ++k) {
        int64_t r = base + (int64_t)k * kBlk;
        if (r < P.D) { double2 a = P.w[r]; s += a.x * a.x + a.y * a.y; }
    }
    double2 t = block_sum2(s, 0.0, sh);
    if (threadIdx.x == 0) P.partial[lt] = t;
}

__global__ void __launch_bounds__(kBlk)
k_bnorm_reduce(const BatchVecDesc *__restrict__ descs)
{
    __shared__ double2 sh[kBlk / 64];
    const BatchVecDesc &P = descs[blockIdx.x];
    double ar = 0.0, ai = 0.0;
    for (int64_t b = threadIdx.x; b < P.ntiles; b += kBlk) { ar += P.partial[b].x; ai += P.partial[b].y; }
    double2 s = block_sum2(ar, ai, sh);
    if (threadIdx.x == 0) P.beta[0] = sqrt(s.x);
}

__global__ void __launch_bounds__(kBlk)
k_bscale(const BatchVecDesc *__restrict__ descs, const int *__restrict__ blk2prob)
{
    const BatchVecDesc &P = descs[blk2prob[blockIdx.x]];
    int64_t r = ((int64_t)blockIdx.x - P.blk0) * kBlk + threadIdx.x;
    if (r >= P.D) return;
    double inv = 1.0 / P.beta[0];
    double2 a = P.w[r];
    P.w[r] = make_double2(a.x * inv, a.y * inv);
}

__global__ void __launch_bounds__(kBlk)
k_brotate(const BatchRotDesc *__restrict__ jobs, const int *__restrict__ blk2job)
{
    extern __shared__ double ys[];   // m * kRotChunk
    const BatchRotDesc &P = jobs[blk2job[blockIdx.x]];
    for (int t = threadIdx.x; t < P.m * kRotChunk; t += kBlk) {
        int i = t / kRotChunk, c = t % kRotChunk;
        ys[t] = c < P.kc ? P.Y[i + (int64_t)P.ldy * (P.c0 + c)] : 0.0;
    }
    __syncthreads();
    int64_t r = ((int64_t)blockIdx.x - P.blk0) * kBlk + threadIdx.x;
    if (r >= P.D) return;
    double ar[kRotChunk], ai[kRotChunk];
#pragma unroll
    for (int c = 0; c < kRotChunk; ++c) { ar[c] = 0.0; ai[c] = 0.0; }
    for (int i = 0; i < P.m; ++i) {
        double2 v = P.V[i][r];
#pragma unroll
        for (int c = 0; c < kRotChunk; ++c) {
            double y = ys[i * kRotChunk + c];
            ar[c] += v.x * y;
            ai[c] += v.y * y;
        }
    }
    for (int c = 0; c < P.kc; ++c) P.out[P.c0 + c][r] = make_double2(ar[c], ai[c]);
}

static_assert(kBlk == kBatchRowBlock && kTile == kBatchTile && kRotChunk == kBatchRotChunk,
              "batched launch geometry shared with the host engine");

inline unsigned rows_blocks(int64_t D) { return (unsigned)((D + kBlk - 1) / kBlk); }
inline unsigned tile_blocks(int64_t D) { return (unsigned)((D + kTile - 1) / kTile); }

}  // namespace

void blas_fill_random(edm_ctx *ctx, double2 *v, int64_t D, uint64_t seed, int mode)
{
    if (D == 0) return;
    hipLaunchKernelGGL(k_fill_random, dim3(rows_blocks(D)), dim3(kBlk), 0, ctx->stream, v, D,
                       seed, mode);
    EDM_HIP_CHECK(hipGetLastError());
}

void blas_multidot(edm_ctx *ctx, const double2 *const *d_vptrs, int nvec, const double2 *w,
                   int64_t D, double2 *d_h)
{
    EDM_REQUIRE(nvec <= kMaxVec, EDM_EINVAL, "too many vectors in one multi-dot");
    if (nvec == 0) return;
    unsigned nb = std::max(1u, tile_blocks(D));
    double2 *part = (double2 *)ctx->partials((size_t)nb * nvec * sizeof(double2));
    const int span = dot_span(nb, nvec);
    const unsigned ny = (unsigned)((nvec + span - 1) / span);
    hipLaunchKernelGGL(k_multidot_partial, dim3(nb, ny), dim3(kBlk), 0, ctx->stream, d_vptrs,
                       nvec, span, w, D, part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(nvec), dim3(kBlk), 0, ctx->stream, part,
                       (int64_t)nb, d_h);
    EDM_HIP_CHECK(hipGetLastError());
}

void blas_multiaxpy(edm_ctx *ctx, const double2 *const *d_vptrs, int nvec, const double2 *d_h,
                    double2 *w, int64_t D)
{
    EDM_REQUIRE(nvec <= kMaxVec, EDM_EINVAL, "too many vectors in one multi-axpy");
    if (nvec == 0 || D == 0) return;
    hipLaunchKernelGGL(k_multiaxpy, dim3(rows_blocks(D)), dim3(kBlk), 0, ctx->stream, d_vptrs,
                       nvec, d_h, w, D);
    EDM_HIP_CHECK(hipGetLastError());
}

void blas_cgs2(edm_ctx *ctx, const double2 *const *d_vptrs, int nvec, double2 *w, int64_t D,
               double2 *d_h, double2 *d_h2)
{
    blas_multidot(ctx, d_vptrs, nvec, w, D, d_h);
    blas_multiaxpy(ctx, d_vptrs, nvec, d_h, w, D);
    blas_multidot(ctx, d_vptrs, nvec, w, D, d_h2);
    blas_multiaxpy(ctx, d_vptrs, nvec, d_h2, w, D);
}

void blas_norm(edm_ctx *ctx, const double2 *w, int64_t D, double *d_out)
{
    unsigned nb = std::max(1u, tile_blocks(D));
    double2 *part = (double2 *)ctx->partials((size_t)(nb + 1) * sizeof(double2));
    hipLaunchKernelGGL(k_norm2_partial, dim3(nb), dim3(kBlk), 0, ctx->stream, w, D, part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlk), 0, ctx->stream, part, (int64_t)nb,
                       part + nb);
    hipLaunchKernelGGL(k_sqrt_first, dim3(1), dim3(64), 0, ctx->stream, part + nb, d_out);
    EDM_HIP_CHECK(hipGetLastError());
}

void blas_scale_inv(edm_ctx *ctx, double2 *v, const double2 *w, const double *d_beta, int64_t D)
{
    if (D == 0) return;
    hipLaunchKernelGGL(k_scale_inv, dim3(rows_blocks(D)), dim3(kBlk), 0, ctx->stream, v, w,
                       d_beta, D);
    EDM_HIP_CHECK(hipGetLastError());
}

void bblas_cgs2_norm(edm_ctx *ctx, const BatchVecDesc *d_descs, int nprob, int maxvec,
                     const int *d_tile2prob, int64_t ntiles, const int *d_blk2prob, int64_t nblk)
{
    EDM_REQUIRE(maxvec <= kMaxVec, EDM_EINVAL, "too many vectors in one batched multi-dot");
    if (nprob == 0) return;
    hipStream_t st = ctx->stream;
    const int span = dot_span(ntiles, std::max(maxvec, 1));
    const unsigned ny = (unsigned)((std::max(maxvec, 1) + span - 1) / span);
    // The multi-axpy walks its blocks in reverse: it starts on the vectors the multi-dot before
    // it read last, and the multi-dot after it starts where it ended, so each launch begins on
    // data still in the MALL (the running problems' vectors are GBs per launch, the MALL
    // 256 MB). Each block's work is unchanged, so the results are too.
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_bmultidot_partial, dim3((unsigned)ntiles, ny), dim3(kBlk), 0, st,
                           d_descs, d_tile2prob, span);
        hipLaunchKernelGGL(k_breduce, dim3((unsigned)nprob, (unsigned)maxvec), dim3(kBlk), 0, st,
                           d_descs, pass);
        hipLaunchKernelGGL(k_baxpy, dim3((unsigned)nblk), dim3(kBlk), 0, st, d_descs, d_blk2prob,
                           pass, 1);
    }
    hipLaunchKernelGGL(k_bnorm_partial, dim3((unsigned)ntiles), dim3(kBlk), 0, st, d_descs,
                       d_tile2prob);
    hipLaunchKernelGGL(k_bnorm_reduce, dim3((unsigned)nprob), dim3(kBlk), 0, st, d_descs);
    EDM_HIP_CHECK(hipGetLastError());
}

void bblas_scale(edm_ctx *ctx, const BatchVecDesc *d_descs, const int *d_blk2prob, int64_t nblk)
{
    if (nblk == 0) return;
    hipLaunchKernelGGL(k_bscale, dim3((unsigned)nblk), dim3(kBlk), 0, ctx->stream, d_descs,
                       d_blk2prob);
    EDM_HIP_CHECK(hipGetLastError());
}

void bblas_rotate(edm_ctx *ctx, const BatchRotDesc *d_jobs, const int *d_blk2job, int64_t nblk,
                  int maxm)
{
    if (nblk == 0) return;
    size_t sh = (size_t)maxm * kRotChunk * sizeof(double);
    EDM_REQUIRE(sh <= 64 * 1024, EDM_EINVAL, "Krylov dimension too large for the rotation kernel");
    hipLaunchKernelGGL(k_brotate, dim3((unsigned)nblk), dim3(kBlk), sh, ctx->stream, d_jobs,
                       d_blk2job);
    EDM_HIP_CHECK(hipGetLastError());
}

void blas_rotate(edm_ctx *ctx, const double2 *const *d_vptrs, int m, const double *d_Y, int ldy,
                 int k, double2 *const *d_outptrs, int64_t D)
{
    if (D == 0 || k == 0) return;
    size_t sh = (size_t)m * kRotChunk * sizeof(double);
    EDM_REQUIRE(sh <= 64 * 1024, EDM_EINVAL, "Krylov dimension too large for the rotation kernel");
    for (int c0 = 0; c0 < k; c0 += kRotChunk) {
        int kc = std::min(kRotChunk, k - c0);
        hipLaunchKernelGGL(k_rotate, dim3(rows_blocks(D)), dim3(kBlk), sh, ctx->stream, d_vptrs,
                           m, d_Y, ldy, c0, kc, d_outptrs, D);
    }
    EDM_HIP_CHECK(hipGetLastError());
}

void blas_rotate_c(edm_ctx *ctx, const double2 *const *d_vptrs, int m, const double2 *d_Y, int ldy,
                   int k, double2 *const *d_outptrs, int64_t D)
{
    if (D == 0 || k == 0) return;
    size_t sh = (size_t)m * kRotChunk * sizeof(double2);
    EDM_REQUIRE(sh <= 64 * 1024, EDM_EINVAL, "Krylov dimension too large for the rotation kernel");
    for (int c0 = 0; c0 < k; c0 += kRotChunk) {
        int kc = std::min(kRotChunk, k - c0);
        hipLaunchKernelGGL(k_rotate_c, dim3(rows_blocks(D)), dim3(kBlk), sh, ctx->stream, d_vptrs,
                           m, d_Y, ldy, c0, kc, d_outptrs, D);
    }
    EDM_HIP_CHECK(hipGetLastError());
}
