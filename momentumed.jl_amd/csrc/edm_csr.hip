// edm_csr.hip — the stored-matrix path: the Hamiltonian of a sector as a device CSR matrix.
//
// Counterpart of the reference's :sparse method on a GPU: SparseHmltMatrix builds the
// matrix on the host (MomentumED/src/method/sparse_matrix.jl:65-191, one column per thread,
// duplicate entries summed), create_gpu_matrix copies it to the device
// (MomentumED/ext/MomentumEDCUDAExt.jl:147-152) and krylov_matrix_solve runs KrylovKit with
// it. Here the matrix is built on the device from the same term tables as the matrix-free
// map: one pass counts each row's contributions, one writes them as sortable 64-bit keys
// (column, term, sign), a segmented radix sort orders every row by (column, term), and a
// merge pass sums the entries of equal columns in term order. Row i holds <i|H|j>, so
// y = A x equals the matrix-free map's y = H x. Storage: int64 row offsets, int32 column
// indices (ascending within a row), complex128 values: 20 bytes per stored entry.
//
// SpMV: G lanes per row (G from the mean row length), each lane sums every G-th entry of the
// row in order and the lanes combine by a fixed butterfly, so y is deterministic.
#include "edm_internal.h"

#include <hipcub/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <numeric>

namespace {

constexpr int kBlk = 256;

__global__ void __launch_bounds__(kBlk)
k_csr_count_unique(int64_t D, const int64_t *__restrict__ rp, const uint64_t *__restrict__ keys,
                   int64_t *__restrict__ cnt)
{
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    int64_t n = 0;
    uint64_t prev = ~0ULL;
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        uint64_t c = keys[k] >> 32;
        n += c != prev;
        prev = c;
    }
    cnt[r] = n;
}

__global__ void __launch_bounds__(kBlk)
k_csr_merge(int64_t D, const int64_t *__restrict__ rp, const uint64_t *__restrict__ keys,
            const double2 *__restrict__ amp, const int64_t *__restrict__ orp,
            int32_t *__restrict__ cols, double2 *__restrict__ vals)
{
    int64_t r = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (r >= D) return;
    int64_t o = orp[r] - 1;
    uint64_t prev = ~0ULL;
    double vr = 0.0, vi = 0.0;
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        uint64_t key = keys[k];
        uint64_t c = key >> 32;
        if (c != prev) {
            if (o >= orp[r]) vals[o] = make_double2(vr, vi);
            ++o;
            cols[o] = (int32_t)c;
            vr = vi = 0.0;
            prev = c;
        }
        double2 a = amp[(key >> 1) & 0x7FFFFFFFULL];
        if (key & 1ULL) { vr -= a.x; vi -= a.y; } else { vr += a.x; vi += a.y; }
    }
    if (o >= orp[r]) vals[o] = make_double2(vr, vi);
}

template <int G>
__global__ void __launch_bounds__(kBlk)
k_csr_spmv(int64_t D, const int64_t *__restrict__ rp, const int32_t *__restrict__ cols,
           const double2 *__restrict__ vals, const double2 *__restrict__ x,
           double2 *__restrict__ y)
{
    int64_t t = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    int64_t r = t / G;
    int l = (int)(t % G);
    double yr = 0.0, yi = 0.0;
    if (r < D) {
        for (int64_t k = rp[r] + l; k < rp[r + 1]; k += G) {
            double2 a = vals[k];
            double2 xv = x[cols[k]];
            yr += a.x * xv.x - a.y * xv.y;
            yi += a.x * xv.y + a.y * xv.x;
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        yr += __shfl_xor(yr, off, G);
        yi += __shfl_xor(yi, off, G);
    }
    if (r < D && l == 0) y[r] = make_double2(yr, yi);
}

// dense row-major copy: workgroup r zeroes row r of A, then writes the row's stored entries
// (columns are distinct after merging)
__global__ void __launch_bounds__(kBlk)
k_csr_dense(int64_t D, const int64_t *__restrict__ rp, const int32_t *__restrict__ cols,
            const double2 *__restrict__ vals, double2 *__restrict__ A)
{
    const int64_t r = blockIdx.x;
    double2 *row = A + r * D;
    for (int64_t c = threadIdx.x; c < D; c += kBlk) row[c] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int64_t k = rp[r] + threadIdx.x; k < rp[r + 1]; k += kBlk) row[cols[k]] = vals[k];
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlk - 1) / kBlk); }

// exclusive prefix of counts[0..n) into out[0..n] (host; the counts are copied back)
void host_scan(edm_ctx *ctx, const int64_t *d_counts, int64_t n, int64_t *d_out, int64_t *total)
{
    std::vector<int64_t> h((size_t)n + 1, 0);
    if (n > 0)
        EDM_HIP_CHECK(hipMemcpyAsync(h.data() + 1, d_counts, n * sizeof(int64_t),
                                     hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < n; ++i) h[i + 1] += h[i];
    *total = h[n];
    EDM_HIP_CHECK(hipMemcpyAsync(d_out, h.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                                 ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace

edm_csr *csr_build(edm_ctx *ctx, const edm_op *op, const edm_basis *b)
{
    EDM_REQUIRE(op->nbits == b->nbits, EDM_EINVAL, "operator and basis have different bits");
    EDM_REQUIRE(op->fast, EDM_EINVAL,
                "the stored matrix supports one- and two-body number-conserving operators");
    EDM_REQUIRE(op->S_full < (1LL << 31), EDM_EINVAL, "too many terms for the stored matrix");
    EDM_REQUIRE(b->D < (1LL << 31), EDM_EINVAL, "the stored matrix needs D < 2^31");
    const int64_t D = b->D;
    hipStream_t st = ctx->stream;
    edm_csr *m = new edm_csr();
    m->ctx = ctx;
    m->D = D;
    int64_t *d_cnt = nullptr, *d_rp = nullptr;
    uint64_t *d_k0 = nullptr, *d_k1 = nullptr;
    void *d_tmp = nullptr;
    size_t tmp_bytes = 0;
    int64_t raw = 0;
    size_t nb = (size_t)std::max<int64_t>(D, 1) * sizeof(int64_t);
    size_t rb = (size_t)(D + 1) * sizeof(int64_t);
    size_t kb = 0;
    try {
        m->d_rowptr = (int64_t *)ctx->alloc(rb);
        d_cnt = (int64_t *)ctx->alloc(nb);
        d_rp = (int64_t *)ctx->alloc(rb);
        // 1. contributions per row, 2. their keys, 3. rows sorted by (column, term)
        launch_connectivity(ctx, op, b, 0, D, d_cnt, nullptr, nullptr, nullptr, nullptr);
        host_scan(ctx, d_cnt, D, d_rp, &raw);
        EDM_REQUIRE(raw < (1LL << 31), EDM_EINVAL,
                    "more than 2^31 stored entries; use the matrix-free map for this sector");
        kb = (size_t)std::max<int64_t>(raw, 1) * sizeof(uint64_t);
        d_k0 = (uint64_t *)ctx->alloc(kb);
        d_k1 = (uint64_t *)ctx->alloc(kb);
        launch_csr_keys(ctx, op, b, d_rp, d_k0);
        if (raw > 0) {
            EDM_HIP_CHECK(hipcub::DeviceSegmentedRadixSort::SortKeys(
                nullptr, tmp_bytes, d_k0, d_k1, (int)raw, (int)D, d_rp, d_rp + 1, 0, 64, st));
            d_tmp = ctx->alloc(std::max<size_t>(tmp_bytes, 16));
            EDM_HIP_CHECK(hipcub::DeviceSegmentedRadixSort::SortKeys(
                d_tmp, tmp_bytes, d_k0, d_k1, (int)raw, (int)D, d_rp, d_rp + 1, 0, 64, st));
        }
        // 4. distinct columns per row, 5. merged entries
        hipLaunchKernelGGL(k_csr_count_unique, dim3(blocks(D)), dim3(kBlk), 0, st, D, d_rp, d_k1,
                           d_cnt);
        EDM_HIP_CHECK(hipGetLastError());
        host_scan(ctx, d_cnt, D, m->d_rowptr, &m->nnz);
        m->d_cols = (int32_t *)ctx->alloc((size_t)std::max<int64_t>(m->nnz, 1) * sizeof(int32_t));
        m->d_vals = (double2 *)ctx->alloc((size_t)std::max<int64_t>(m->nnz, 1) * sizeof(double2));
        hipLaunchKernelGGL(k_csr_merge, dim3(blocks(D)), dim3(kBlk), 0, st, D, d_rp, d_k1,
                           op->d_amp, m->d_rowptr, m->d_cols, m->d_vals);
        EDM_HIP_CHECK(hipGetLastError());
        EDM_HIP_CHECK(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        if (d_tmp) ctx->free(d_tmp, std::max<size_t>(tmp_bytes, 16));
        if (d_k0) ctx->free(d_k0, kb);
        if (d_k1) ctx->free(d_k1, kb);
        if (d_rp) ctx->free(d_rp, rb);
        if (d_cnt) ctx->free(d_cnt, nb);
        csr_destroy(m);
        throw;
    }
    if (d_tmp) ctx->free(d_tmp, std::max<size_t>(tmp_bytes, 16));
    ctx->free(d_k0, kb);
    ctx->free(d_k1, kb);
    ctx->free(d_rp, rb);
    ctx->free(d_cnt, nb);
    return m;
}

void csr_destroy(edm_csr *m)
{
    if (!m) return;
    edm_ctx *ctx = m->ctx;
    if (m->d_rowptr) ctx->free(m->d_rowptr, (size_t)(m->D + 1) * sizeof(int64_t));
    if (m->d_cols) ctx->free(m->d_cols, (size_t)std::max<int64_t>(m->nnz, 1) * sizeof(int32_t));
    if (m->d_vals) ctx->free(m->d_vals, (size_t)std::max<int64_t>(m->nnz, 1) * sizeof(double2));
    delete m;
}

void csr_matvec(edm_ctx *ctx, const edm_csr *m, const double2 *x, double2 *y)
{
    const int64_t D = m->D;
    if (D == 0) return;
    const double mean = (double)m->nnz / (double)D;
    if (mean < 12)
        hipLaunchKernelGGL(k_csr_spmv<4>, dim3(blocks(D * 4)), dim3(kBlk), 0, ctx->stream, D,
                           m->d_rowptr, m->d_cols, m->d_vals, x, y);
    else if (mean < 48)
        hipLaunchKernelGGL(k_csr_spmv<8>, dim3(blocks(D * 8)), dim3(kBlk), 0, ctx->stream, D,
                           m->d_rowptr, m->d_cols, m->d_vals, x, y);
    else
        hipLaunchKernelGGL(k_csr_spmv<16>, dim3(blocks(D * 16)), dim3(kBlk), 0, ctx->stream, D,
                           m->d_rowptr, m->d_cols, m->d_vals, x, y);
    EDM_HIP_CHECK(hipGetLastError());
}

void csr_dense(edm_ctx *ctx, const edm_csr *m, double2 *A)
{
    if (m->D == 0) return;
    EDM_REQUIRE(m->D <= (1LL << 31) - 1, EDM_EINVAL, "dimension too large for a dense matrix");
    hipLaunchKernelGGL(k_csr_dense, dim3((unsigned)m->D), dim3(kBlk), 0, ctx->stream, m->D,
                       m->d_rowptr, m->d_cols, m->d_vals, A);
    EDM_HIP_CHECK(hipGetLastError());
}
