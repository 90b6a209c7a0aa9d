// edm_api.cpp — extern "C" entry points of libedm.so (declared in include/edm.h).
#include "edm_internal.h"

#include <new>

#include <complex>

#include <algorithm>
#include <cstring>


namespace {

template <class F>
int guarded(edm_ctx *ctx, F &&f)
{
    try {
        int dev = ctx ? ctx->device : -1;
        if (ctx) {
            DeviceGuard g(dev);
            return f();
        }
        return f();
    } catch (const EdmError &e) {
        if (ctx) ctx->last_error = e.msg;
        edm_set_global_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        edm_set_global_error(e.what());
        return EDM_EINVAL;
    }
}

void free_basis(edm_basis *b)
{
    if (!b) return;
    edm_ctx *ctx = b->ctx;
    ctx->free(b->d_states, (size_t)std::max<int64_t>(b->D, 1) * sizeof(uint64_t));
    if (b->d_jlo) ctx->free(b->d_jlo, (1ULL << b->nlo) * sizeof(uint32_t));
    if (b->d_jhi) ctx->free(b->d_jhi, (1ULL << b->nhi) * sizeof(uint32_t));
    if (b->d_jlo16) ctx->free(b->d_jlo16, (1ULL << b->nlo) * sizeof(uint16_t));
    free_narrow_tables(b);
    if (b->d_start) ctx->free(b->d_start, (b->nbuckets + 1) * sizeof(uint32_t));
    delete b;
}

void drop_tables(edm_basis *b)
{
    if (b->d_jlo) b->ctx->free(b->d_jlo, (1ULL << b->nlo) * sizeof(uint32_t));
    if (b->d_jhi) b->ctx->free(b->d_jhi, (1ULL << b->nhi) * sizeof(uint32_t));
    if (b->d_jlo16) b->ctx->free(b->d_jlo16, (1ULL << b->nlo) * sizeof(uint16_t));
    b->d_jlo = b->d_jhi = nullptr;
    b->d_jlo16 = nullptr;
    free_narrow_tables(b);
    b->rank_mode = 0;
    b->rank_derived = false;
}

// keep freshly built rank tables only when they reproduce every state's position
void verify_or_drop(edm_basis *b)
{
    if (verify_rank_tables(b)) {
        b->rank_mode = 1;
        // the LDS low-table split is an optimisation: a budget that cannot hold it keeps the
        // tables above
        try {
            build_narrow_tables(b);
        } catch (const EdmError &e) {
            if (e.code != EDM_ENOMEM) throw;
            free_narrow_tables(b);
        } catch (const std::bad_alloc &) {   // host tables of the two-level split
            free_narrow_tables(b);
        }
    } else {
        drop_tables(b);
    }
}

// the binary search's bucket index for a basis whose lookups may search (no system rank
// tables); an optimisation, so a budget that cannot hold it leaves the plain search
void maybe_index(edm_basis *b)
{
    if (b->rank_mode == 1 && !b->rank_derived) return;
    try {
        build_search_index(b);
    } catch (const EdmError &e) {
        if (e.code != EDM_ENOMEM) throw;
    }
}

void attach_rank(edm_basis *b)
{
    b->rank_mode = 0;
    if (build_rank_tables(b)) verify_or_drop(b);
}

}  // namespace

#define EDM_NULL_BATCH(bt)                                   \
    do {                                                     \
        if (!(bt)) {                                         \
            edm_set_global_error("batch handle is NULL");    \
            return EDM_EINVAL;                               \
        }                                                    \
    } while (0)

extern "C" {

int edm_device_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return 0;
    }
    return 1;
}

int edm_ctx_create(int device, uint64_t hbm_budget_bytes, edm_ctx **out)
{
    return guarded(nullptr, [&]() -> int {
        EDM_REQUIRE(out != nullptr, EDM_EINVAL, "out is NULL");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            throw EdmError{EDM_ENODEV,
                           "device=:rocm requires a functional AMD GPU; none is visible"};
        }
        EDM_REQUIRE(device >= 0 && device < n, EDM_ENODEV, "no such GPU index");
        DeviceGuard g(device);
        edm_ctx *c = new edm_ctx();
        c->device = device;
        c->budget = hbm_budget_bytes;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw EdmError{EDM_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
        }
        try {
            c->d_noalloc = noalloc_scratch_address();
        } catch (...) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *out = c;
        return EDM_OK;
    });
}

int edm_ctx_destroy(edm_ctx *ctx)
{
    if (!ctx) return EDM_OK;
    {
        DeviceGuard g(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->d_binom) ctx->free(ctx->d_binom, 65 * 65 * sizeof(uint64_t));
        if (ctx->d_partials) ctx->free(ctx->d_partials, ctx->partials_bytes);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return EDM_OK;
}

const char *edm_last_error(const edm_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : edm_global_error();
}

int edm_synchronize(edm_ctx *ctx)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx, EDM_EINVAL, "ctx is NULL");
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return EDM_OK;
    });
}

int edm_ctx_stream(edm_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return EDM_EINVAL;
    *stream = (void *)ctx->stream;
    return EDM_OK;
}

int edm_ctx_memory(edm_ctx *ctx, uint64_t *used, uint64_t *budget)
{
    if (!ctx) return EDM_EINVAL;
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    return EDM_OK;
}

int edm_num_sectors(const edm_system *sys, int64_t *nsec)
{
    return guarded(nullptr, [&]() -> int {
        SystemSpec sp = SystemSpec::from(sys);
        EDM_REQUIRE(nsec, EDM_EINVAL, "nsec is NULL");
        *nsec = sp.nsec;
        return EDM_OK;
    });
}

int edm_sector_sizes(edm_ctx *ctx, const edm_system *sys, int64_t *counts)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && counts, EDM_EINVAL, "NULL argument");
        SystemSpec sp = SystemSpec::from(sys);
        std::vector<int64_t> c;
        enum_sector_counts(ctx, sp, c);
        std::copy(c.begin(), c.end(), counts);
        return EDM_OK;
    });
}

int edm_basis_enumerate(edm_ctx *ctx, const edm_system *sys, int64_t sector, edm_basis **out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && out, EDM_EINVAL, "NULL argument");
        *out = nullptr;
        SystemSpec sp = SystemSpec::from(sys);
        EDM_REQUIRE(sector >= 0 && sector < sp.nsec, EDM_EINVAL, "sector index out of range");
        edm_basis *b = new edm_basis();
        b->ctx = ctx;
        b->nbits = sp.nbits;
        b->has_spec = true;
        b->spec = sp;
        b->sector = sector;
        try {
            b->d_states = enum_sector_states(ctx, sp, sector, &b->D);
            attach_rank(b);
            maybe_index(b);
        } catch (...) {
            free_basis(b);
            throw;
        }
        *out = b;
        return EDM_OK;
    });
}

int edm_basis_enumerate_all(edm_ctx *ctx, const edm_system *sys, edm_basis **out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && sys && out, EDM_EINVAL, "NULL argument");
        SystemSpec sp = SystemSpec::from(sys);
        for (int64_t k = 0; k < sp.nsec; ++k) out[k] = nullptr;
        std::vector<uint64_t *> states;
        std::vector<int64_t> D;
        enum_all_sector_states(ctx, sp, states, D);
        try {
            std::vector<edm_basis *> bs;
            for (int64_t k = 0; k < sp.nsec; ++k) {
                edm_basis *b = new edm_basis();
                b->ctx = ctx;
                b->nbits = sp.nbits;
                b->has_spec = true;
                b->spec = sp;
                b->sector = k;
                b->D = D[k];
                b->d_states = states[k];
                states[k] = nullptr;
                out[k] = b;
                bs.push_back(b);
            }
            std::vector<bool> built = build_rank_tables_many(bs);
            for (size_t k = 0; k < bs.size(); ++k) {
                bs[k]->rank_mode = 0;
                if (built[k]) verify_or_drop(bs[k]);
                maybe_index(bs[k]);
            }
        } catch (...) {
            for (int64_t q = 0; q < sp.nsec; ++q) {
                if (out[q]) free_basis(out[q]);
                out[q] = nullptr;
                if (states[q]) ctx->free(states[q], (size_t)std::max<int64_t>(D[q], 1) * sizeof(uint64_t));
            }
            throw;
        }
        return EDM_OK;
    });
}

int edm_basis_upload(edm_ctx *ctx, const uint64_t *states, int64_t D, int32_t nbits,
                     const edm_system *sys, int64_t sector, edm_basis **out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && out, EDM_EINVAL, "NULL argument");
        *out = nullptr;
        EDM_REQUIRE(D >= 0 && (D == 0 || states), EDM_EINVAL, "bad basis array");
        EDM_REQUIRE(nbits >= 0 && nbits <= 64, EDM_EINVAL, "nbits must be 0..64");
        uint64_t lim = nbits >= 64 ? ~0ULL : ((1ULL << nbits) - 1ULL);
        for (int64_t i = 0; i < D; ++i) {
            EDM_REQUIRE((states[i] & ~lim) == 0, EDM_EINVAL,
                        "State representation out of range for given bits");
            if (i) EDM_REQUIRE(states[i - 1] < states[i], EDM_EINVAL,
                               "basis must be sorted ascending without repeats");
        }
        edm_basis *b = new edm_basis();
        b->ctx = ctx;
        b->D = D;
        b->nbits = nbits;
        try {
            if (sys) {
                b->spec = SystemSpec::from(sys);
                EDM_REQUIRE(b->spec.nbits == nbits, EDM_EINVAL, "nbits differs from the system's");
                b->has_spec = true;
                b->sector = sector;
            }
            b->d_states = (uint64_t *)ctx->alloc((size_t)std::max<int64_t>(D, 1) * sizeof(uint64_t));
            if (D)
                EDM_HIP_CHECK(hipMemcpyAsync(b->d_states, states, D * sizeof(uint64_t),
                                             hipMemcpyHostToDevice, ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            if (sys && sector >= 0) attach_rank(b);
            if (b->rank_mode == 0) {
                // derived tables are an optimisation: an upload that cannot hold them (HBM
                // budget) still succeeds, with binary search
                bool got = false;
                try {
                    got = derive_rank_tables(b, states);
                } catch (const EdmError &) {
                    got = false;
                }
                if (got) {
                    b->rank_derived = true;
                    verify_or_drop(b);
                    if (b->rank_mode == 0) b->rank_derived = false;
                } else {
                    drop_tables(b);
                }
            }
            maybe_index(b);
        } catch (...) {
            free_basis(b);
            throw;
        }
        *out = b;
        return EDM_OK;
    });
}

int edm_basis_size(const edm_basis *b, int64_t *D)
{
    if (!b || !D) return EDM_EINVAL;
    *D = b->D;
    return EDM_OK;
}

int edm_basis_download(const edm_basis *b, uint64_t *states_host)
{
    if (!b) return EDM_EINVAL;
    return guarded(b->ctx, [&]() -> int {
        EDM_REQUIRE(states_host || b->D == 0, EDM_EINVAL, "NULL output");
        if (b->D)
            EDM_HIP_CHECK(hipMemcpyAsync(states_host, b->d_states, b->D * sizeof(uint64_t),
                                         hipMemcpyDeviceToHost, b->ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        return EDM_OK;
    });
}

int edm_basis_device_ptr(const edm_basis *b, const uint64_t **states_dev)
{
    if (!b || !states_dev) return EDM_EINVAL;
    *states_dev = b->d_states;
    return EDM_OK;
}

int edm_basis_rank_mode(const edm_basis *b, int32_t *mode)
{
    if (!b || !mode) return EDM_EINVAL;
    *mode = b->rank_mode == 1 && b->rank_derived ? 2 : b->rank_mode;
    return EDM_OK;
}

int edm_basis_drop_rank(edm_basis *b)
{
    if (!b) return EDM_EINVAL;
    return guarded(b->ctx, [&]() -> int {
        EDM_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        drop_tables(b);
        maybe_index(b);
        return EDM_OK;
    });
}

int edm_basis_destroy(edm_basis *b)
{
    if (!b) return EDM_OK;
    return guarded(b->ctx, [&]() -> int {
        (void)hipStreamSynchronize(b->ctx->stream);
        free_basis(b);
        return EDM_OK;
    });
}

int edm_op_upload(edm_ctx *ctx, int32_t nbits, const double *amp, const uint64_t *out,
                  const uint64_t *in, int64_t S, int32_t upper_hermitian, edm_op **op)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op, EDM_EINVAL, "NULL argument");
        *op = build_op(ctx, nbits, amp, out, in, S, upper_hermitian != 0);
        return EDM_OK;
    });
}

int edm_op_size(const edm_op *op, int64_t *S_full)
{
    if (!op || !S_full) return EDM_EINVAL;
    *S_full = op->S_full;
    return EDM_OK;
}

int edm_op_full_list(const edm_op *op, double *amp, uint64_t *out, uint64_t *in, uint64_t *parity)
{
    if (!op) return EDM_EINVAL;
    size_t n = (size_t)op->S_full;
    if (amp) std::memcpy(amp, op->amp.data(), 2 * n * sizeof(double));
    if (out) std::memcpy(out, op->out.data(), n * sizeof(uint64_t));
    if (in) std::memcpy(in, op->in.data(), n * sizeof(uint64_t));
    if (parity) std::memcpy(parity, op->par.data(), n * sizeof(uint64_t));
    return EDM_OK;
}

int edm_op_destroy(edm_op *op)
{
    if (!op) return EDM_OK;
    return guarded(op->ctx, [&]() -> int {
        (void)hipStreamSynchronize(op->ctx->stream);
        destroy_op(op);
        return EDM_OK;
    });
}

int edm_matvec(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const void *x_dev, void *y_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && x_dev && y_dev, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(x_dev != y_dev, EDM_EINVAL, "x and y must not alias");
        launch_matvec(ctx, op, b, (const double2 *)x_dev, (double2 *)y_dev, ApplyMode::Forward);
        return EDM_OK;
    });
}

int edm_matvec_adjoint(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const void *x_dev,
                       void *y_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && x_dev && y_dev, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(x_dev != y_dev, EDM_EINVAL, "x and y must not alias");
        launch_matvec(ctx, op, b, (const double2 *)x_dev, (double2 *)y_dev, ApplyMode::Adjoint);
        return EDM_OK;
    });
}

int edm_matvec_block(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int nv,
                     const void *const *x_dev, void *const *y_dev, void *work_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && x_dev && y_dev && work_dev, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(nv >= 1 && nv <= 64, EDM_EINVAL, "nv must be in 1..64");
        for (int v = 0; v < nv; ++v) {
            EDM_REQUIRE(x_dev[v] && y_dev[v], EDM_EINVAL, "NULL vector pointer");
            for (int u = 0; u < nv; ++u)
                EDM_REQUIRE(x_dev[u] != y_dev[v], EDM_EINVAL, "x and y must not alias");
            EDM_REQUIRE(x_dev[v] != work_dev && y_dev[v] != work_dev, EDM_EINVAL,
                        "work must not alias x or y");
        }
        launch_matvec_block(ctx, op, b, nv, (const double2 *const *)x_dev, (double2 *)work_dev,
                            (double2 *const *)y_dev);
        return EDM_OK;
    });
}

int edm_matvec_host(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const double *x, double *y)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && x && y, EDM_EINVAL, "NULL argument");
        size_t bytes = (size_t)std::max<int64_t>(b->D, 1) * sizeof(double2);
        double2 *dx = (double2 *)ctx->alloc(bytes);
        double2 *dy = nullptr;
        try {
            dy = (double2 *)ctx->alloc(bytes);
            EDM_HIP_CHECK(hipMemcpyAsync(dx, x, b->D * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
            launch_matvec(ctx, op, b, dx, dy, ApplyMode::Forward);
            EDM_HIP_CHECK(hipMemcpyAsync(y, dy, b->D * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->free(dx, bytes);
            ctx->free(dy, bytes);
            throw;
        }
        ctx->free(dx, bytes);
        ctx->free(dy, bytes);
        return EDM_OK;
    });
}

int edm_connectivity_counts(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t row0,
                            int64_t nrows, int64_t *counts)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && counts, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(row0 >= 0 && nrows >= 0 && row0 + nrows <= b->D, EDM_EINVAL, "row range");
        size_t bytes = (size_t)std::max<int64_t>(nrows, 1) * sizeof(int64_t);
        int64_t *d = (int64_t *)ctx->alloc(bytes);
        launch_connectivity(ctx, op, b, row0, nrows, d, nullptr, nullptr, nullptr, nullptr);
        EDM_HIP_CHECK(hipMemcpyAsync(counts, d, nrows * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->free(d, bytes);
        return EDM_OK;
    });
}

int edm_connectivity(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t row0,
                     int64_t nrows, const int64_t *rowptr, int64_t *cols, int32_t *terms,
                     int8_t *negs)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && rowptr && cols && terms && negs, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(row0 >= 0 && nrows >= 0 && row0 + nrows <= b->D, EDM_EINVAL, "row range");
        int64_t nnz = rowptr[nrows];
        // the fill kernel writes rowptr[r]..rowptr[r+1]-1 for row r: check the caller's
        // offsets against freshly computed counts before any device write
        {
            std::vector<int64_t> cnt((size_t)std::max<int64_t>(nrows, 1));
            size_t cbytes = cnt.size() * sizeof(int64_t);
            int64_t *d_cnt = (int64_t *)ctx->alloc(cbytes);
            launch_connectivity(ctx, op, b, row0, nrows, d_cnt, nullptr, nullptr, nullptr, nullptr);
            EDM_HIP_CHECK(hipMemcpyAsync(cnt.data(), d_cnt, nrows * sizeof(int64_t),
                                         hipMemcpyDeviceToHost, ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            ctx->free(d_cnt, cbytes);
            EDM_REQUIRE(rowptr[0] == 0, EDM_EINVAL, "rowptr[0] must be 0");
            for (int64_t r = 0; r < nrows; ++r)
                EDM_REQUIRE(rowptr[r + 1] - rowptr[r] == cnt[r], EDM_EINVAL,
                            "rowptr is not the exclusive prefix of edm_connectivity_counts");
        }
        size_t rb = (size_t)(nrows + 1) * sizeof(int64_t);
        size_t cb = (size_t)std::max<int64_t>(nnz, 1) * sizeof(int64_t);
        size_t tb = (size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t);
        size_t nb = (size_t)std::max<int64_t>(nnz, 1) * sizeof(int8_t);
        int64_t *d_rp = (int64_t *)ctx->alloc(rb);
        int64_t *d_c = (int64_t *)ctx->alloc(cb);
        int32_t *d_t = (int32_t *)ctx->alloc(tb);
        int8_t *d_n = (int8_t *)ctx->alloc(nb);
        EDM_HIP_CHECK(hipMemcpyAsync(d_rp, rowptr, rb, hipMemcpyHostToDevice, ctx->stream));
        launch_connectivity(ctx, op, b, row0, nrows, nullptr, d_rp, d_c, d_t, d_n);
        EDM_HIP_CHECK(hipMemcpyAsync(cols, d_c, nnz * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        EDM_HIP_CHECK(hipMemcpyAsync(terms, d_t, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        EDM_HIP_CHECK(hipMemcpyAsync(negs, d_n, nnz * sizeof(int8_t), hipMemcpyDeviceToHost, ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->free(d_rp, rb);
        ctx->free(d_c, cb);
        ctx->free(d_t, tb);
        ctx->free(d_n, nb);
        // rows sorted by (column, term)
        std::vector<int64_t> idx;
        std::vector<int64_t> c2;
        std::vector<int32_t> t2;
        std::vector<int8_t> n2;
        for (int64_t r = 0; r < nrows; ++r) {
            int64_t a = rowptr[r], e = rowptr[r + 1];
            idx.resize(e - a);
            for (int64_t k = 0; k < e - a; ++k) idx[k] = a + k;
            std::sort(idx.begin(), idx.end(), [&](int64_t p, int64_t q) {
                return cols[p] < cols[q] || (cols[p] == cols[q] && terms[p] < terms[q]);
            });
            c2.resize(e - a); t2.resize(e - a); n2.resize(e - a);
            for (int64_t k = 0; k < e - a; ++k) {
                c2[k] = cols[idx[k]]; t2[k] = terms[idx[k]]; n2[k] = negs[idx[k]];
            }
            for (int64_t k = 0; k < e - a; ++k) {
                cols[a + k] = c2[k]; terms[a + k] = t2[k]; negs[a + k] = n2[k];
            }
        }
        return EDM_OK;
    });
}

int edm_connectivity_totals(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t *nnz,
                            int64_t *nnz_offdiag)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && nnz && nnz_offdiag, EDM_EINVAL, "NULL argument");
        launch_conn_totals(ctx, op, b, nnz, nnz_offdiag);
        return EDM_OK;
    });
}

int edm_eigsolve(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t nev,
                 int32_t krylovdim, double tol, int32_t maxiter, uint64_t seed,
                 const double *vec0, double *vals, double *vecs, edm_solve_info *info)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && vals, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(op->nbits == b->nbits, EDM_EDIM,
                    "Dimension of Hamiltonian linear map mismatches vector length.");
        int rc = run_eigsolve(ctx, op, b, nev, krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
        if (rc == EDM_ENOCONV) {
            ctx->last_error = "Krylov method fails. Cannot find " + std::to_string(nev) +
                              " eigenvectors within maxiter.";
        }
        return rc;
    });
}

int edm_eigsolve_general(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t nev,
                         int32_t krylovdim, double tol, int32_t maxiter, uint64_t seed,
                         const double *vec0, double *vals, double *vecs, edm_solve_info *info)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && vals, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(op->nbits == b->nbits, EDM_EDIM,
                    "Dimension of Hamiltonian linear map mismatches vector length.");
        int rc = run_eigsolve_general_mv(
            ctx, b->D,
            [&](const double2 *x, double2 *y) { launch_matvec(ctx, op, b, x, y, ApplyMode::Forward); },
            nev, krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
        if (rc == EDM_ENOCONV) {
            ctx->last_error = "Krylov method fails. Cannot find " + std::to_string(nev) +
                              " eigenvectors within maxiter.";
        }
        return rc;
    });
}

int edm_eigsolve_block(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t nev,
                       int32_t block, int32_t krylovdim, double tol, int32_t maxiter,
                       uint64_t seed, const double *vec0, double *vals, double *vecs,
                       edm_solve_info *info)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && vals, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(op->nbits == b->nbits, EDM_EDIM,
                    "Dimension of Hamiltonian linear map mismatches vector length.");
        int rc = run_eigsolve_block_op(ctx, op, b, block, nev, krylovdim, tol, maxiter, seed, vec0,
                                       vals, vecs, info);
        if (rc == EDM_ENOCONV) {
            ctx->last_error = "Krylov method fails. Cannot find " + std::to_string(nev) +
                              " eigenvectors within maxiter.";
        }
        return rc;
    });
}

int edm_fill_random(edm_ctx *ctx, void *v_dev, int64_t D, uint64_t seed, int32_t mode)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && (v_dev || D == 0), EDM_EINVAL, "NULL argument");
        blas_fill_random(ctx, (double2 *)v_dev, D, seed, mode);
        return EDM_OK;
    });
}

int edm_csr_build(edm_ctx *ctx, const edm_op *op, const edm_basis *b, edm_csr **out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && out, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL,
                    "operator and basis must belong to the context");
        *out = nullptr;
        *out = csr_build(ctx, op, b);
        return EDM_OK;
    });
}

int edm_csr_size(const edm_csr *m, int64_t *D, int64_t *nnz)
{
    return guarded(m ? m->ctx : nullptr, [&]() -> int {
        EDM_REQUIRE(m, EDM_EINVAL, "NULL matrix");
        if (D) *D = m->D;
        if (nnz) *nnz = m->nnz;
        return EDM_OK;
    });
}

int edm_csr_download(const edm_csr *m, int64_t *rowptr, int32_t *cols, double *vals)
{
    return guarded(m ? m->ctx : nullptr, [&]() -> int {
        EDM_REQUIRE(m && rowptr && cols && vals, EDM_EINVAL, "NULL argument");
        hipStream_t st = m->ctx->stream;
        EDM_HIP_CHECK(hipMemcpyAsync(rowptr, m->d_rowptr, (m->D + 1) * sizeof(int64_t),
                                     hipMemcpyDeviceToHost, st));
        if (m->nnz > 0) {
            EDM_HIP_CHECK(hipMemcpyAsync(cols, m->d_cols, m->nnz * sizeof(int32_t),
                                         hipMemcpyDeviceToHost, st));
            EDM_HIP_CHECK(hipMemcpyAsync(vals, m->d_vals, m->nnz * sizeof(double2),
                                         hipMemcpyDeviceToHost, st));
        }
        EDM_HIP_CHECK(hipStreamSynchronize(st));
        return EDM_OK;
    });
}

int edm_csr_matvec(edm_ctx *ctx, const edm_csr *m, const void *x_dev, void *y_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && m && (m->D == 0 || (x_dev && y_dev)), EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(m->ctx == ctx, EDM_EINVAL, "matrix belongs to another context");
        EDM_REQUIRE(x_dev != y_dev || m->D == 0, EDM_EINVAL, "x and y must not alias");
        csr_matvec(ctx, m, (const double2 *)x_dev, (double2 *)y_dev);
        return EDM_OK;
    });
}

int edm_csr_eigsolve(edm_ctx *ctx, const edm_csr *m, int32_t nev, int32_t krylovdim, double tol,
                     int32_t maxiter, uint64_t seed, const double *vec0, double *vals,
                     double *vecs, edm_solve_info *info)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && m && vals, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(m->ctx == ctx, EDM_EINVAL, "matrix belongs to another context");
        int rc = run_eigsolve_mv(
            ctx, m->D, [&](const double2 *x, double2 *y) { csr_matvec(ctx, m, x, y); }, nev,
            krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
        if (rc == EDM_ENOCONV) ctx->last_error = "Krylov method fails. Cannot find N eigenvectors.";
        return rc;
    });
}

int edm_csr_eigsolve_general(edm_ctx *ctx, const edm_csr *m, int32_t nev, int32_t krylovdim,
                             double tol, int32_t maxiter, uint64_t seed, const double *vec0,
                             double *vals, double *vecs, edm_solve_info *info)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && m && vals, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(m->ctx == ctx, EDM_EINVAL, "matrix belongs to another context");
        int rc = run_eigsolve_general_mv(
            ctx, m->D, [&](const double2 *x, double2 *y) { csr_matvec(ctx, m, x, y); }, nev,
            krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
        if (rc == EDM_ENOCONV) ctx->last_error = "Krylov method fails. Cannot find N eigenvectors.";
        return rc;
    });
}

int edm_csr_dense(edm_ctx *ctx, const edm_csr *m, void *a_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && m && a_dev, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(m->ctx == ctx, EDM_EINVAL, "matrix belongs to another context");
        csr_dense(ctx, m, (double2 *)a_dev);
        return EDM_OK;
    });
}

int edm_csr_destroy(edm_csr *m)
{
    if (!m) return EDM_OK;
    return guarded(m->ctx, [&]() -> int {
        (void)hipStreamSynchronize(m->ctx->stream);
        csr_destroy(m);
        return EDM_OK;
    });
}

int edm_matvec_tune(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t *kernel,
                    double *ms)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && kernel, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        double t[EDM_TUNE_CANDIDATES];
        *kernel = matvec_tune(ctx, op, b, t);
        if (ms)
            for (int k = 0; k < EDM_TUNE_CANDIDATES; ++k) ms[k] = t[k];
        return EDM_OK;
    });
}

int edm_matvec_kernel(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t *kernel)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && b && kernel, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL, "handles from another context");
        *kernel = matvec_kernel_code(op, b);
        return EDM_OK;
    });
}

int edm_matvec_between(edm_ctx *ctx, const edm_op *op, const edm_basis *bra,
                       const edm_basis *ket, const void *x_dev, void *y_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op && bra && ket, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op->ctx == ctx && bra->ctx == ctx && ket->ctx == ctx, EDM_EINVAL,
                    "handles from another context");
        EDM_REQUIRE((bra->D == 0 || y_dev) && (ket->D == 0 || x_dev), EDM_EINVAL, "NULL vector");
        EDM_REQUIRE(x_dev != y_dev || bra->D == 0, EDM_EINVAL, "x and y must not alias");
        launch_matvec_between(ctx, op, bra, ket, (const double2 *)x_dev, (double2 *)y_dev);
        return EDM_OK;
    });
}

int edm_rdm1(edm_ctx *ctx, const edm_basis *b, const void *psi_dev, int32_t npairs,
             const int32_t *pairs, double *out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && b && out && (npairs == 0 || pairs), EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(b->ctx == ctx, EDM_EINVAL, "basis from another context");
        EDM_REQUIRE(b->D == 0 || psi_dev, EDM_EINVAL, "NULL vector");
        EDM_REQUIRE(npairs >= 0 && npairs <= 65535, EDM_EINVAL, "npairs must be 0..65535");
        std::fill(out, out + 2 * (size_t)npairs, 0.0);
        if (b->D == 0 || npairs == 0) return EDM_OK;
        // a complete sector's rank tables hold every target of a sector-conserving c_i^+ c_j;
        // every target of any other pair lies outside the sector (entry exactly 0)
        const bool lin = b->rank_mode == 1 && b->has_spec && !b->rank_derived;
        std::vector<int2> use;
        std::vector<int> where;
        for (int p = 0; p < npairs; ++p) {
            int i = pairs[2 * p], j = pairs[2 * p + 1];
            EDM_REQUIRE(i >= 0 && j >= 0 && i < b->nbits && j < b->nbits, EDM_EINVAL,
                        "orbital index outside the basis bits");
            if (lin && i != j) {
                const SystemSpec &sp = b->spec;
                bool keep = sp.comp_of_bit(i) == sp.comp_of_bit(j);
                for (int d = 0; keep && d < sp.dim; ++d) {
                    int64_t dk = sp.orb_k(i, d) - sp.orb_k(j, d);
                    keep = sp.gk[d] != 0 ? dk % sp.gk[d] == 0 : dk == 0;
                }
                if (!keep) continue;
            }
            use.push_back(make_int2(i, j));
            where.push_back(p);
        }
        if (use.empty()) return EDM_OK;
        const int n = (int)use.size();
        size_t pb = n * sizeof(int2), ob = n * sizeof(double2);
        int2 *d_pairs = (int2 *)ctx->alloc(pb);
        double2 *d_out = (double2 *)ctx->alloc(ob);
        std::vector<double2> h(n);
        try {
            EDM_HIP_CHECK(hipMemcpyAsync(d_pairs, use.data(), pb, hipMemcpyHostToDevice, ctx->stream));
            launch_rdm1(ctx, b, (const double2 *)psi_dev, d_pairs, n, lin, d_out);
            EDM_HIP_CHECK(hipMemcpyAsync(h.data(), d_out, ob, hipMemcpyDeviceToHost, ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->free(d_pairs, pb);
            ctx->free(d_out, ob);
            throw;
        }
        ctx->free(d_pairs, pb);
        ctx->free(d_out, ob);
        for (int q = 0; q < n; ++q) {
            out[2 * where[q]] = h[q].x;
            out[2 * where[q] + 1] = h[q].y;
        }
        return EDM_OK;
    });
}

int edm_bracket2(edm_ctx *ctx, const edm_op *op2, const edm_op *op1, const edm_basis *bra,
                 const edm_basis *ket, const void *bra_dev, const void *ket_dev, double *out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op2 && op1 && bra && ket && out, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op2->ctx == ctx && op1->ctx == ctx && bra->ctx == ctx && ket->ctx == ctx,
                    EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(op1->nbits == ket->nbits && op2->nbits == bra->nbits && bra->nbits == ket->nbits,
                    EDM_EDIM, "operators and bases have different bits");
        out[0] = out[1] = 0.0;
        if (ket->D == 0 || bra->D == 0) return EDM_OK;
        EDM_REQUIRE(bra_dev && ket_dev, EDM_EINVAL, "NULL vector");
        double2 *d_out = (double2 *)ctx->alloc(sizeof(double2));
        double2 h = make_double2(0.0, 0.0);
        try {
            launch_bracket2(ctx, op2, op1, bra, ket, (const double2 *)bra_dev,
                            (const double2 *)ket_dev, d_out);
            EDM_HIP_CHECK(hipMemcpyAsync(&h, d_out, sizeof(double2), hipMemcpyDeviceToHost,
                                         ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->free(d_out, sizeof(double2));
            throw;
        }
        ctx->free(d_out, sizeof(double2));
        out[0] = h.x;
        out[1] = h.y;
        return EDM_OK;
    });
}

int edm_bracket2_noalloc(edm_ctx *ctx, const edm_op *op2, const edm_op *op1,
                         const edm_basis *bra, const edm_basis *ket, const void *bra_dev,
                         const void *ket_dev, void *out_dev)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && op2 && op1 && bra && ket && out_dev, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(op2->ctx == ctx && op1->ctx == ctx && bra->ctx == ctx && ket->ctx == ctx,
                    EDM_EINVAL, "handles from another context");
        EDM_REQUIRE(op1->nbits == ket->nbits && op2->nbits == bra->nbits && bra->nbits == ket->nbits,
                    EDM_EDIM, "operators and bases have different bits");
        EDM_REQUIRE((ket->D == 0 || ket_dev) && (bra->D == 0 || bra_dev), EDM_EINVAL,
                    "NULL vector");
        launch_bracket2_noalloc(ctx, op2, op1, bra, ket, (const double2 *)bra_dev,
                                (const double2 *)ket_dev, (double2 *)out_dev);
        return EDM_OK;
    });
}

int edm_pes_block_matrix(edm_ctx *ctx, const edm_basis *psi_basis, const void *psi_dev,
                         int64_t nA, const uint64_t *A, uint64_t ph_mask, const edm_basis *B,
                         void *m_dev, int64_t *counts)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && psi_basis && B && (A || nA == 0), EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(psi_basis->ctx == ctx && B->ctx == ctx, EDM_EINVAL,
                    "handles from another context");
        EDM_REQUIRE(nA >= 0, EDM_EINVAL, "nA must be >= 0");
        if (counts) counts[0] = counts[1] = 0;
        if (nA == 0 || psi_basis->D == 0) return EDM_OK;
        EDM_REQUIRE(psi_dev && (m_dev || B->D == 0), EDM_EINVAL, "NULL vector");
        EDM_REQUIRE(nA <= INT64_MAX / psi_basis->D, EDM_EINVAL, "D * nA overflows int64");
        const size_t ab = (size_t)nA * sizeof(uint64_t), cb = 2 * sizeof(unsigned long long);
        uint64_t *d_A = (uint64_t *)ctx->alloc(ab);
        unsigned long long *d_counts = nullptr;
        unsigned long long h[2] = {0, 0};
        try {
            d_counts = (unsigned long long *)ctx->alloc(cb);
            EDM_HIP_CHECK(hipMemcpyAsync(d_A, A, ab, hipMemcpyHostToDevice, ctx->stream));
            EDM_HIP_CHECK(hipMemsetAsync(d_counts, 0, cb, ctx->stream));
            launch_pes_fill(ctx, psi_basis, (const double2 *)psi_dev, nA, d_A, ph_mask, B,
                            (double2 *)m_dev, d_counts);
            EDM_HIP_CHECK(hipMemcpyAsync(h, d_counts, cb, hipMemcpyDeviceToHost, ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->free(d_A, ab);
            if (d_counts) ctx->free(d_counts, cb);
            throw;
        }
        ctx->free(d_A, ab);
        ctx->free(d_counts, cb);
        if (counts) {
            counts[0] = (int64_t)h[0];
            counts[1] = (int64_t)h[1];
            return EDM_OK;
        }
        EDM_REQUIRE(h[0] == h[1], EDM_EINVAL,
                    "a complement state is missing from the block B (psi's subspace must hold "
                    "one momentum sector and B the complementary block)");
        return EDM_OK;
    });
}

int edm_step_inner_prod(edm_ctx *ctx, const edm_basis *b, int32_t g, const void *const *vf_dev,
                        const void *const *vi_dev, const double *orbital_ip, double *out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && b && vf_dev && vi_dev && orbital_ip && out, EDM_EINVAL, "NULL argument");
        EDM_REQUIRE(b->ctx == ctx, EDM_EINVAL, "basis from another context");
        EDM_REQUIRE(g >= 1 && g <= 64, EDM_EINVAL, "g must be 1..64");
        std::fill(out, out + 2 * (size_t)g * g, 0.0);
        if (b->D == 0) return EDM_OK;
        size_t ub = std::max(b->nbits, 1) * sizeof(double2), tb = 2 * (size_t)g * sizeof(void *);
        size_t ob = (size_t)g * g * sizeof(double2);
        double2 *d_u = (double2 *)ctx->alloc(ub);
        const double2 **d_tab = (const double2 **)ctx->alloc(tb);
        double2 *d_out = (double2 *)ctx->alloc(ob);
        std::vector<const void *> tab(2 * (size_t)g);
        for (int a = 0; a < g; ++a) {
            EDM_REQUIRE(vf_dev[a] && vi_dev[a], EDM_EINVAL, "NULL vector");
            tab[a] = vf_dev[a];
            tab[g + a] = vi_dev[a];
        }
        std::vector<double2> h((size_t)g * g);
        try {
            if (b->nbits > 0)
                EDM_HIP_CHECK(hipMemcpyAsync(d_u, orbital_ip, b->nbits * sizeof(double2),
                                             hipMemcpyHostToDevice, ctx->stream));
            EDM_HIP_CHECK(hipMemcpyAsync(d_tab, tab.data(), tb, hipMemcpyHostToDevice, ctx->stream));
            launch_step_gram(ctx, b, d_u, d_tab, d_tab + g, g, d_out);
            EDM_HIP_CHECK(hipMemcpyAsync(h.data(), d_out, ob, hipMemcpyDeviceToHost, ctx->stream));
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->free(d_u, ub);
            ctx->free(d_tab, tb);
            ctx->free(d_out, ob);
            throw;
        }
        ctx->free(d_u, ub);
        ctx->free(d_tab, tb);
        ctx->free(d_out, ob);
        for (size_t q = 0; q < h.size(); ++q) {
            out[2 * q] = h[q].x;
            out[2 * q + 1] = h[q].y;
        }
        return EDM_OK;
    });
}

int edm_projected_eigh(int32_t n, const double *A, double *w, double *Z)
{
    return guarded(nullptr, [&]() -> int {
        EDM_REQUIRE(n >= 1 && A && w && Z, EDM_EINVAL, "bad argument");
        std::vector<double> a(A, A + (size_t)n * n), ww, zz;
        sym_eigh(n, a, ww, zz);
        std::copy(ww.begin(), ww.end(), w);
        std::copy(zz.begin(), zz.end(), Z);
        return EDM_OK;
    });
}

int edm_projected_heigh(int32_t n, const double *A, double *w, double *Z)
{
    return guarded(nullptr, [&]() -> int {
        EDM_REQUIRE(n >= 1 && A && w && Z, EDM_EINVAL, "bad argument");
        std::vector<std::complex<double>> a((size_t)n * n), zz;
        for (size_t q = 0; q < a.size(); ++q) a[q] = std::complex<double>(A[2 * q], A[2 * q + 1]);
        std::vector<double> ww;
        herm_eigh(n, a, ww, zz);
        std::copy(ww.begin(), ww.end(), w);
        for (size_t q = 0; q < zz.size(); ++q) {
            Z[2 * q] = zz[q].real();
            Z[2 * q + 1] = zz[q].imag();
        }
        return EDM_OK;
    });
}

int edm_projected_schur(int32_t n, const double *A, double *T, double *Z)
{
    return guarded(nullptr, [&]() -> int {
        EDM_REQUIRE(n >= 1 && A && T && Z, EDM_EINVAL, "bad argument");
        return projected_schur(n, A, T, Z);
    });
}

int edm_batch_create(edm_ctx *ctx, edm_batch **out)
{
    return guarded(ctx, [&]() -> int {
        EDM_REQUIRE(ctx && out, EDM_EINVAL, "NULL argument");
        *out = batch_create(ctx);
        return EDM_OK;
    });
}

int edm_batch_destroy(edm_batch *bt)
{
    if (!bt) return EDM_OK;
    return guarded(batch_ctx(bt), [&]() -> int {
        batch_destroy(bt);
        return EDM_OK;
    });
}

int edm_batch_add(edm_batch *bt, const edm_op *op, const edm_basis *b, int32_t nev,
                  int32_t krylovdim, double tol, int32_t maxiter, uint64_t seed,
                  const double *vec0, int32_t want_vectors, int64_t *id)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int {
        return batch_add(bt, op, b, nev, krylovdim, tol, maxiter, seed, vec0, want_vectors, id);
    });
}

int edm_batch_step(edm_batch *bt, int32_t max_steps, int32_t *n_running, int32_t *n_finished)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int {
        return batch_step(bt, max_steps, n_running, n_finished);
    });
}

int edm_batch_finished(edm_batch *bt, int64_t *ids, int32_t cap, int32_t *n)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int {
        EDM_REQUIRE(n && cap >= 0, EDM_EINVAL, "bad argument");
        return batch_finished(bt, ids, cap, n);
    });
}

int edm_batch_status(edm_batch *bt, int64_t id, int32_t *state, edm_solve_info *info)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int { return batch_status(bt, id, state, info); });
}

int edm_batch_result(edm_batch *bt, int64_t id, double *vals, double *vecs, edm_solve_info *info)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int { return batch_result(bt, id, vals, vecs, info); });
}

int edm_batch_remove(edm_batch *bt, int64_t id)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int { return batch_remove(bt, id); });
}

int edm_batch_count(edm_batch *bt, int32_t *n_total, int32_t *n_running)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int { return batch_count(bt, n_total, n_running); });
}

int edm_batch_set_quantum(edm_batch *bt, int64_t rows)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int { return batch_set_quantum(bt, rows); });
}

int edm_batch_set_weight(edm_batch *bt, int64_t id, double weight)
{
    EDM_NULL_BATCH(bt);
    return guarded(batch_ctx(bt), [&]() -> int { return batch_set_weight(bt, id, weight); });
}

int edm_batch_problem_bytes(int64_t D, int32_t nev, int32_t krylovdim, uint64_t *bytes)
{
    return guarded(nullptr, [&]() -> int {
        EDM_REQUIRE(bytes && D > 0 && nev >= 1, EDM_EINVAL, "bad argument");
        *bytes = batch_problem_bytes(D, nev, krylovdim);
        return EDM_OK;
    });
}

}  // extern "C"
