// edm_batch.cpp — many Lanczos eigensolves advanced together on one device.
//
// The reference solves one (sector, parameter) problem at a time: EDsolve -> krylov_map_solve
// -> KrylovKit.eigsolve (MomentumED/src/MomentumED.jl:375-480, method/linear_map.jl:179-191),
// and a sweep is a host loop over sectors and parameter points. Small sectors leave a GPU
// idle between the dozen launches and two host round trips of every Lanczos step. Here a
// batch holds any number of independent problems, each with its own basis, term table and
// Krylov vectors resident in HBM. One step advances every running problem by one Lanczos
// iteration with one launch per kernel for the whole batch (matvec, two Gram-Schmidt passes,
// norm, scale), one host-to-device plan upload and one device-to-host read of all projection
// coefficients. Restarts and final Ritz-vector rotations of all problems that end a cycle in
// the same step share one launch as well.
//
// Every problem follows exactly the iteration of run_eigsolve (edm_lanczos.cpp): the same
// start vector, the same restart rule, the same breakdown handling, and kernels whose
// per-problem arithmetic and reduction order equal the single-problem kernels'. A problem's
// eigenvalues and eigenvectors are therefore bitwise the ones edm_eigsolve returns for it
// alone, whatever else shares the batch. Problems may be added and collected between steps
// (continuous batching); each add takes the problem's workspace from the context (with an
// HBM budget, one allocation per problem, so the budget is the admission limit; without,
// a range of a shared 1 GiB chunk).
#include "edm_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <atomic>
#include <climits>
#include <condition_variable>
#include <exception>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

namespace {

constexpr size_t kAlign = 256;

inline size_t align_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

int64_t tiles_of(int64_t D) { return std::max<int64_t>(1, (D + kBatchTile - 1) / kBatchTile); }
int64_t rblocks_of(int64_t D) { return (D + kBatchRowBlock - 1) / kBatchRowBlock; }

// Pinned host staging that grows on demand. Callers only rewrite it after the stream has
// passed every copy that reads it.
struct PinnedBuf {
    unsigned char *p = nullptr;
    size_t bytes = 0;
    void ensure(hipStream_t st, size_t need)
    {
        if (need <= bytes) return;
        EDM_HIP_CHECK(hipStreamSynchronize(st));
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        size_t n = align_up(need + need / 2 + kAlign);
        EDM_HIP_CHECK(hipHostMalloc((void **)&p, n, hipHostMallocDefault));
        bytes = n;
    }
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

struct DevArena {
    void *p = nullptr;
    size_t bytes = 0;
    void ensure(edm_ctx *ctx, size_t need)
    {
        if (need <= bytes) return;
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (p) ctx->free(p, bytes);
        p = nullptr;
        bytes = 0;
        size_t n = align_up(need + need / 2 + kAlign);
        p = ctx->alloc(n);
        bytes = n;
    }
    void release(edm_ctx *ctx)
    {
        if (p) ctx->free(p, bytes);
        p = nullptr;
        bytes = 0;
    }
};

// Problem workspaces. A context with an HBM budget gives each problem its own allocation, so
// the budget admits exactly what fits. Without a budget, workspaces are carved first-fit
// from 1 GiB chunks and returned to the chunk's free list when the problem is collected; a
// chunk goes back to the device allocator once nothing in it is in use. Collecting thousands
// of small problems then costs no driver call each (a device free of a workspace of a few MB
// takes ~0.3 ms). All work on a workspace is queued on the context's one stream, so a range
// handed to a new problem is only written after the work of its previous owner.
constexpr size_t kSlabChunk = (size_t)1 << 30;

struct SlabArena {
    struct Chunk {
        unsigned char *base;
        size_t size, in_use;
        std::map<size_t, size_t> free;   // offset -> length, non-adjacent ranges
    };
    std::vector<Chunk> chunks;

    void *alloc(edm_ctx *ctx, size_t bytes)
    {
        bytes = align_up(bytes);
        for (Chunk &c : chunks)
            for (auto it = c.free.begin(); it != c.free.end(); ++it) {
                if (it->second < bytes) continue;
                const size_t off = it->first, len = it->second;
                c.free.erase(it);
                if (len > bytes) c.free[off + bytes] = len - bytes;
                c.in_use += bytes;
                return c.base + off;
            }
        const size_t sz = ctx->budget ? bytes : std::max(bytes, kSlabChunk);
        Chunk c{(unsigned char *)ctx->alloc(sz), sz, bytes, {}};
        if (sz > bytes) c.free[bytes] = sz - bytes;
        chunks.push_back(std::move(c));
        return chunks.back().base;
    }

    void free(edm_ctx *ctx, void *p, size_t bytes)
    {
        bytes = align_up(bytes);
        unsigned char *q = (unsigned char *)p;
        for (size_t k = 0; k < chunks.size(); ++k) {
            Chunk &c = chunks[k];
            if (q < c.base || q >= c.base + c.size) continue;
            c.in_use -= bytes;
            if (c.in_use == 0) {
                ctx->free(c.base, c.size);
                chunks.erase(chunks.begin() + (ptrdiff_t)k);
                return;
            }
            size_t off = (size_t)(q - c.base), len = bytes;
            auto next = c.free.lower_bound(off);
            if (next != c.free.end() && next->first == off + len) {
                len += next->second;
                next = c.free.erase(next);
            }
            if (next != c.free.begin()) {
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    prev->second += len;
                    return;
                }
            }
            c.free[off] = len;
            return;
        }
    }

    void release_all(edm_ctx *ctx)
    {
        for (Chunk &c : chunks) ctx->free(c.base, c.size);
        chunks.clear();
    }
};

// Host builder for one upload: sections appended at 256-byte alignment.
struct Plan {
    std::vector<unsigned char> h;
    void clear() { h.clear(); }
    size_t put(const void *src, size_t n)
    {
        size_t off = align_up(h.size());
        h.resize(off + n);
        if (n) memcpy(h.data() + off, src, n);
        return off;
    }
    size_t reserve(size_t n)
    {
        size_t off = align_up(h.size());
        h.resize(off + n);
        return off;
    }
};

// problem states reported by edm_batch_status (include/edm.h)
enum ProblemState { kRunning = 0, kDone = 1, kNotConverged = 2 };

// A fixed set of host threads for independent per-problem work. run(n, f) calls f(i) for
// every i < n (the calling thread takes part) and returns when all calls have returned.
class HostPool {
public:
    HostPool()
    {
        unsigned hc = std::thread::hardware_concurrency();
        int nw = (int)std::min(16u, std::max(1u, hc)) - 1;
        for (int w = 0; w < nw; ++w) workers_.emplace_back([this] { loop(); });
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }
    void run(int n, const std::function<void(int)> &f)
    {
        if (n <= 0) return;
        if (n == 1 || workers_.empty()) {
            for (int i = 0; i < n; ++i) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &f;
            n_ = n;
            next_.store(0);
            active_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return active_ == 0; });
        job_ = nullptr;
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    void work()
    {
        for (;;) {
            int i = next_.fetch_add(1);
            if (i >= n_) return;
            try {
                (*job_)(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu_);
                if (!err_) err_ = std::current_exception();
            }
        }
    }
    void loop()
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            {
                std::lock_guard<std::mutex> g(mu_);
                if (--active_ == 0) done_cv_.notify_all();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *job_ = nullptr;
    int n_ = 0;
    std::atomic<int> next_{0};
    int active_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
};

}  // namespace

struct edm_problem {
    int64_t id = 0;
    const edm_op *op = nullptr;
    const edm_basis *b = nullptr;
    int64_t D = 0, Dpad = 0, ntiles = 0;
    int nev = 0, m = 0, maxiter = 0;
    double tol = 0.0;
    uint64_t seed = 0;
    bool want_vectors = false;
    // device workspace: one slab
    void *slab = nullptr;
    size_t slab_bytes = 0;
    std::vector<double2 *> V, tmp;   // m + 1 Krylov vectors, m rotation targets
    double2 *partial = nullptr;      // [m + 1][ntiles]
    double2 *d_h = nullptr, *d_h2 = nullptr;
    double *d_beta = nullptr, *d_Y = nullptr;
    double2 **d_vptrs = nullptr;     // for the single-problem kernels (start, breakdown)
    // Lanczos state (as in run_eigsolve)
    std::vector<double> T, theta, Y, res;
    int k = 0, j = 0, iters = 1, matvecs = 0, nconv = 0;
    double beta_last = 0.0, hnorm = 0.0;
    uint64_t rng_stream = 0;
    int state = kRunning;
    int64_t credit = 0;              // rows of step quantum accumulated (set_quantum)
    double weight = 1.0;             // credit per step = weight * quantum (set_weight)
    std::vector<double> vals;
    std::chrono::steady_clock::time_point t0;
    double seconds = 0.0;
    double &t(int a, int c) { return T[(size_t)a * m + c]; }
};

// One interleave job of a step: the G x vectors (D rows) of a shared-walk group copied into xi.
struct XJob {
    int G;
    int64_t D;
    const double2 *x[8];
    double2 *xi;
};

// Rows from which a shared-walk group gathers x from an interleaved copy of its vectors: below,
// the vectors are small enough for the L2 and the copy would not pay. EDM_BATCH_INTERLEAVE_MIN
// overrides (0: every group, -1: none).
// Groups of this many rows or more get their own launch of 1,024-row workgroups, whose LDS
// can hold the 14-bit rank split (the shared 512-row launch keeps three workgroups per CU, too
// little LDS for it): 48-orbital sector, 8 ratios batched, 24.0 -> 21.1 s. EDM_BATCH_BIG_GROUP=0
// keeps them in the shared launch; EDM_BATCH_BIG_GROUP_ROWS=n moves the threshold.
static int64_t big_group_rows()
{
    const char *v = getenv("EDM_BATCH_BIG_GROUP_ROWS");
    return v ? std::max<int64_t>(0, (int64_t)atoll(v)) : (int64_t)1 << 18;
}
static bool batch_big_groups()
{
    const char *v = getenv("EDM_BATCH_BIG_GROUP");
    return !(v && atoi(v) == 0);
}

static int64_t batch_interleave_min()
{
    const char *v = getenv("EDM_BATCH_INTERLEAVE_MIN");
    if (!v) return (int64_t)1 << 18;
    return (int64_t)atoll(v);
}

struct edm_batch {
    edm_ctx *ctx = nullptr;
    std::map<int64_t, edm_problem *> probs;   // ordered by id: launch order is admission order
    PinnedBuf h_plan, h_coef, h_rot;
    DevArena d_plan, d_coef, d_rot;
    DevArena d_xi;               // interleaved x copies of large shared-walk groups (per step)
    std::vector<XJob> xjobs;     // this step's interleave jobs
    SlabArena slabs;
    HostPool pool;
    // host scratch of the per-step plan, kept between steps so that building a plan does
    // not allocate
    Plan plan;
    std::vector<unsigned char> mvdesc;
    std::vector<int> mvblk, tile2prob, blk2prob;
    // shared-walk matvec descriptors and block tables by group size (index 2..max group)
    std::vector<unsigned char> gdesc[8];
    std::vector<int> gblk[8];
    // the same for groups of big_group_rows() rows and more (own launch, 1,024-row workgroups)
    std::vector<unsigned char> gdesc_big[8];
    std::vector<int> gblk_big[8];
    std::vector<size_t> coef_off, vtab_off;
    std::vector<BatchVecDesc> vd;
    int64_t steps = 0;
    int64_t quantum = 0;   // 0: every running problem takes part in every step
};

static size_t problem_slab_bytes(int64_t D, int m)
{
    int64_t Dpad = (D + 63) / 64 * 64;
    size_t vb = (size_t)Dpad * sizeof(double2);
    size_t b = vb * (size_t)(2 * m + 1);
    b += align_up((size_t)(m + 1) * tiles_of(D) * sizeof(double2));   // partials
    b += align_up((size_t)(m + 1) * sizeof(double2)) * 2;             // h, h2
    b += align_up(2 * sizeof(double));                                // beta
    b += align_up((size_t)m * m * sizeof(double));                    // Y
    b += align_up((size_t)(m + 1) * sizeof(double2 *));               // pointer table
    return b;
}

static int krylov_dim(int64_t D, int nev, int krylovdim)
{
    int m = std::max(krylovdim, nev + 2);
    if ((int64_t)m > D) m = (int)D;
    return m;
}

static void release_problem(edm_batch *bt, edm_problem *p)
{
    if (p->slab) bt->slabs.free(bt->ctx, p->slab, p->slab_bytes);
    p->slab = nullptr;
    delete p;
}

// CGS2 of w against the problem's V[0..n) with the single-problem kernels; as cgs2 in
// edm_lanczos.cpp.
static void single_cgs2(edm_ctx *ctx, edm_problem *p, int n, double2 *w, std::vector<double2> &h)
{
    h.assign(n, make_double2(0.0, 0.0));
    if (n == 0) return;
    EDM_HIP_CHECK(hipMemcpyAsync(p->d_vptrs, p->V.data(), p->V.size() * sizeof(double2 *),
                                 hipMemcpyHostToDevice, ctx->stream));
    std::vector<double2> h2(n);
    blas_cgs2(ctx, p->d_vptrs, n, w, p->D, p->d_h, p->d_h2);
    EDM_HIP_CHECK(hipMemcpyAsync(h.data(), p->d_h, n * sizeof(double2), hipMemcpyDeviceToHost,
                                 ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(h2.data(), p->d_h2, n * sizeof(double2), hipMemcpyDeviceToHost,
                                 ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
        h[i].x += h2[i].x;
        h[i].y += h2[i].y;
    }
}

static double single_norm(edm_ctx *ctx, edm_problem *p, const double2 *v)
{
    blas_norm(ctx, v, p->D, p->d_beta);
    double b = 0.0;
    EDM_HIP_CHECK(hipMemcpyAsync(&b, p->d_beta, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return b;
}

int batch_add(edm_batch *bt, const edm_op *op, const edm_basis *b, int nev, int krylovdim,
              double tol, int maxiter, uint64_t seed, const double *vec0, int want_vectors,
              int64_t *id)
{
    edm_ctx *ctx = bt->ctx;
    EDM_REQUIRE(op && b && id, EDM_EINVAL, "null argument");
    EDM_REQUIRE(op->ctx == ctx && b->ctx == ctx, EDM_EINVAL,
                "operator and basis must belong to the batch's context");
    EDM_REQUIRE(op->nbits == b->nbits, EDM_EINVAL, "operator and basis have different bits");
    const int64_t D = b->D;
    EDM_REQUIRE(D > 0, EDM_EINVAL, "empty Hilbert subspace");
    EDM_REQUIRE(nev >= 1, EDM_EINVAL, "N should be at least 1.");
    EDM_REQUIRE(nev <= D, EDM_EINVAL, "more eigenpairs requested than the subspace dimension");
    EDM_REQUIRE(tol > 0, EDM_EINVAL, "tol must be positive");
    EDM_REQUIRE(maxiter >= 1, EDM_EINVAL, "maxiter must be at least 1");
    int m = krylov_dim(D, nev, krylovdim);
    EDM_REQUIRE(m <= 255, EDM_EINVAL, "krylovdim must be at most 255");
    EDM_REQUIRE(D < ((int64_t)1 << 31), EDM_EINVAL, "batched problems must have D < 2^31");

    edm_problem *p = new edm_problem();
    p->t0 = std::chrono::steady_clock::now();
    p->id = (int64_t)edm_next_id();
    p->op = op;
    p->b = b;
    p->D = D;
    p->Dpad = (D + 63) / 64 * 64;
    p->ntiles = tiles_of(D);
    p->nev = nev;
    p->m = m;
    p->maxiter = maxiter;
    p->tol = tol;
    p->seed = seed;
    p->want_vectors = want_vectors != 0;
    try {
        p->slab_bytes = problem_slab_bytes(D, m);
        p->slab = bt->slabs.alloc(ctx, p->slab_bytes);
        unsigned char *c = (unsigned char *)p->slab;
        size_t vb = (size_t)p->Dpad * sizeof(double2);
        for (int i = 0; i <= m; ++i, c += vb) p->V.push_back((double2 *)c);
        for (int i = 0; i < m; ++i, c += vb) p->tmp.push_back((double2 *)c);
        p->partial = (double2 *)c;
        c += align_up((size_t)(m + 1) * p->ntiles * sizeof(double2));
        p->d_h = (double2 *)c;
        c += align_up((size_t)(m + 1) * sizeof(double2));
        p->d_h2 = (double2 *)c;
        c += align_up((size_t)(m + 1) * sizeof(double2));
        p->d_beta = (double *)c;
        c += align_up(2 * sizeof(double));
        p->d_Y = (double *)c;
        c += align_up((size_t)m * m * sizeof(double));
        p->d_vptrs = (double2 **)c;

        // start vector, as run_eigsolve; a caller's vec0 is checked on the host, the seeded
        // uniform [0,1)+i[0,1) one cannot have a zero or non-finite norm and is normalised
        // on the device without a round trip
        hipStream_t st = ctx->stream;
        if (vec0) {
            EDM_HIP_CHECK(hipMemcpyAsync(p->V[0], vec0, D * sizeof(double2), hipMemcpyHostToDevice, st));
            double b0 = single_norm(ctx, p, p->V[0]);
            EDM_REQUIRE(b0 > 0 && std::isfinite(b0), EDM_EINVAL, "start vector is zero or not finite");
        } else {
            blas_fill_random(ctx, p->V[0], D, seed, 0);
            blas_norm(ctx, p->V[0], D, p->d_beta);
        }
        blas_scale_inv(ctx, p->V[0], p->V[0], p->d_beta, D);
        p->T.assign((size_t)m * m, 0.0);
        p->res.assign(m, 0.0);
        p->k = 0;
        p->j = 0;
        p->iters = 1;
    } catch (...) {
        release_problem(bt, p);
        throw;
    }
    bt->probs[p->id] = p;
    *id = p->id;
    return EDM_OK;
}

// Host side of one Lanczos step of problem p, given h = V^H w (both passes summed) and
// beta = ||w|| after orthogonalisation; w has already been scaled by 1/beta on the device.
// As the inner loop of run_eigsolve.
static void host_step(edm_ctx *ctx, edm_problem *p, std::vector<double2> &h, double beta)
{
    const int j = p->j, m = p->m;
    p->t(j, j) = h[j].x;
    p->hnorm = std::max(p->hnorm, std::fabs(h[j].x) + beta + (j > 0 ? std::fabs(p->t(j, j - 1)) : 0.0));
    bool breakdown = !(beta > 1e-14 * std::max(p->hnorm, 1e-300));
    if (breakdown) {
        beta = 0.0;
        for (int attempt = 0; attempt < 4; ++attempt) {
            blas_fill_random(ctx, p->V[j + 1], p->D, p->seed ^ (0xA5A5A5A5ULL + ++p->rng_stream), 0);
            std::vector<double2> hh;
            single_cgs2(ctx, p, j + 1, p->V[j + 1], hh);
            double nb = single_norm(ctx, p, p->V[j + 1]);
            if (nb > 1e-8) break;
        }
        blas_scale_inv(ctx, p->V[j + 1], p->V[j + 1], p->d_beta, p->D);
    }
    if (j + 1 < m) {
        p->t(j, j + 1) = beta;
        p->t(j + 1, j) = beta;
    }
    p->beta_last = beta;
    p->j = j + 1;
}

// End of a cycle: Rayleigh-Ritz, convergence test; returns the number of rotation columns to
// apply (keep for a restart, nev for the final vectors, 0 when nothing is rotated) and sets
// the problem's state.
static int host_cycle_end(edm_problem *p, bool *finishing)
{
    const int m = p->m;
    std::vector<double> A = p->T;
    sym_eigh(m, A, p->theta, p->Y);
    p->nconv = 0;
    for (int i = 0; i < m; ++i) p->res[i] = std::fabs(p->beta_last * p->Y[(size_t)i * m + (m - 1)]);
    while (p->nconv < p->nev && p->res[p->nconv] <= p->tol) ++p->nconv;
    if (p->nconv >= p->nev || p->iters == p->maxiter || (int64_t)m == p->D) {
        *finishing = true;
        p->vals.assign(p->theta.begin(), p->theta.begin() + p->nev);
        return p->want_vectors ? p->nev : 0;
    }
    *finishing = false;
    int keep = (3 * m + 2 * p->nconv) / 5;
    keep = std::max(keep, p->nev);
    keep = std::min(keep, m - 1);
    return keep;
}

// Host bookkeeping after a restart's rotation was queued (run_eigsolve's restart block).
static void host_restart(edm_problem *p, int keep)
{
    const int m = p->m;
    for (int c = 0; c < keep; ++c) std::swap(p->V[c], p->tmp[c]);
    std::swap(p->V[keep], p->V[m]);
    std::fill(p->T.begin(), p->T.end(), 0.0);
    for (int c = 0; c < keep; ++c) {
        p->t(c, c) = p->theta[c];
        double cpl = p->beta_last * p->Y[(size_t)c * m + (m - 1)];
        p->t(c, keep) = cpl;
        p->t(keep, c) = cpl;
    }
    p->k = keep;
    p->j = keep;
    p->iters += 1;
}

static void finish_problem(edm_problem *p)
{
    p->state = p->nconv >= p->nev ? kDone : kNotConverged;
    p->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - p->t0).count();
}

// Queue the rotations of every (problem, number of columns) in `rots` with one upload and
// one launch: tmp[c] = sum_i V[i] Y[i, c].
static void queue_rotations(edm_batch *bt, const std::vector<std::pair<edm_problem *, int>> &rots)
{
    if (rots.empty()) return;
    edm_ctx *ctx = bt->ctx;
    Plan pl;
    std::vector<BatchRotDesc> jobs;
    std::vector<int> blk2job;
    int maxm = 0;
    // per problem: V table, out table, Y columns
    struct Off { size_t v, o, y; };
    std::vector<Off> offs;
    for (auto &r : rots) {
        edm_problem *p = r.first;
        Off o;
        o.v = pl.put(p->V.data(), (size_t)p->m * sizeof(double2 *));
        o.o = pl.put(p->tmp.data(), (size_t)r.second * sizeof(double2 *));
        o.y = pl.put(p->Y.data(), (size_t)p->m * r.second * sizeof(double));
        offs.push_back(o);
    }
    // descriptors refer to device addresses of the plan sections: filled after sizing the arena
    int64_t nblk = 0;
    for (size_t q = 0; q < rots.size(); ++q) {
        edm_problem *p = rots[q].first;
        int ncol = rots[q].second;
        for (int c0 = 0; c0 < ncol; c0 += kBatchRotChunk) {
            BatchRotDesc d{};
            d.D = p->D;
            d.m = p->m;
            d.ldy = p->m;
            d.c0 = c0;
            d.kc = std::min(kBatchRotChunk, ncol - c0);
            d.blk0 = nblk;
            int64_t nb = rblocks_of(p->D);
            blk2job.insert(blk2job.end(), (size_t)nb, (int)jobs.size());
            nblk += nb;
            jobs.push_back(d);
            maxm = std::max(maxm, p->m);
        }
    }
    size_t jobs_off = pl.reserve(jobs.size() * sizeof(BatchRotDesc));
    size_t blk_off = pl.put(blk2job.data(), blk2job.size() * sizeof(int));
    bt->d_rot.ensure(ctx, pl.h.size());
    unsigned char *dev = (unsigned char *)bt->d_rot.p;
    size_t ji = 0;
    for (size_t q = 0; q < rots.size(); ++q) {
        int ncol = rots[q].second;
        for (int c0 = 0; c0 < ncol; c0 += kBatchRotChunk, ++ji) {
            jobs[ji].V = (const double2 *const *)(dev + offs[q].v);
            jobs[ji].out = (double2 *const *)(dev + offs[q].o);
            jobs[ji].Y = (const double *)(dev + offs[q].y);
        }
    }
    memcpy(pl.h.data() + jobs_off, jobs.data(), jobs.size() * sizeof(BatchRotDesc));
    bt->h_rot.ensure(ctx->stream, pl.h.size());
    memcpy(bt->h_rot.p, pl.h.data(), pl.h.size());
    EDM_HIP_CHECK(hipMemcpyAsync(bt->d_rot.p, bt->h_rot.p, pl.h.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    bblas_rotate(ctx, (const BatchRotDesc *)(dev + jobs_off), (const int *)(dev + blk_off), nblk, maxm);
}

// One Lanczos step of every running problem. Returns the number of problems that finished.
static int batch_step_once(edm_batch *bt)
{
    edm_ctx *ctx = bt->ctx;
    hipStream_t st = ctx->stream;
    // Problems whose last step completed a cycle (j == m) are parked: their Rayleigh-Ritz
    // runs on the host thread pool while this step's device work runs for the others, and
    // their rotation is queued behind it; they take part again in the next step.
    std::vector<edm_problem *> act, parked;
    for (auto &kv : bt->probs) {
        edm_problem *p = kv.second;
        if (p->state != kRunning) continue;
        (p->j == p->m ? parked : act).push_back(p);
    }
    if (act.empty() && parked.empty()) return 0;
    if (bt->quantum > 0 && !act.empty()) {
        // deficit round robin over rows: each step credits every running problem with
        // weight * `quantum` rows; a problem takes part when its credit covers its
        // dimension, so a problem of D rows steps about once every D / (weight * quantum)
        // steps and every problem of weight 1 receives about the same device work per step
        auto per_step = [&](const edm_problem *p) {
            return std::max<int64_t>(1, (int64_t)std::llround(p->weight * (double)bt->quantum));
        };
        int64_t rounds = INT64_MAX;
        for (edm_problem *p : act) {
            const int64_t q = per_step(p);
            rounds = std::min(rounds, std::max<int64_t>(1, (p->D - p->credit + q - 1) / q));
        }
        std::vector<edm_problem *> sel;
        for (edm_problem *p : act) {
            p->credit += rounds * per_step(p);
            if (p->credit >= p->D) {
                p->credit = std::min(p->credit - p->D, p->D);
                sel.push_back(p);
            }
        }
        act.swap(sel);
    }
    const int n = (int)act.size();

    // coefficient arena: per problem h[m+1], h2[m+1], beta (one double2 slot)
    std::vector<size_t> &coef_off = bt->coef_off;
    coef_off.resize(n);
    size_t coef_bytes = 0;
    for (int q = 0; q < n; ++q) {
        coef_off[q] = coef_bytes;
        coef_bytes += (size_t)(2 * (act[q]->m + 1) + 1) * sizeof(double2);
    }
    bt->d_coef.ensure(ctx, coef_bytes);
    bt->h_coef.ensure(st, coef_bytes);

    // plan: matvec descriptors + 512-row block table, V tables, vector descriptors, tile and
    // 256-row block tables
    Plan &pl = bt->plan;
    pl.clear();
    const size_t mvsz = batch_mv_desc_size();
    std::vector<unsigned char> &mvdesc = bt->mvdesc;
    std::vector<int> &mvblk = bt->mvblk;
    mvdesc.clear();
    mvblk.clear();
    std::vector<edm_problem *> single_mv;
    size_t lds = 16;
    // Problems on one basis whose operators share the term structure (the sweep's parameter
    // points of one sector) go through one row walk, up to batch_mv_max_group() at a time;
    // every other problem gets its own descriptor.
    const int maxg = std::min(batch_mv_max_group(), 7);
    const size_t gsz = batch_mv_group_desc_size();
    for (int g = 2; g <= maxg; ++g) {
        bt->gdesc[g].clear();
        bt->gblk[g].clear();
        bt->gdesc_big[g].clear();
        bt->gblk_big[g].clear();
    }
    size_t lds_big[8] = {0};
    const bool big_groups = batch_big_groups();
    const int64_t big_rows = big_group_rows();
    std::vector<edm_problem *> solo;
    {
        std::map<std::pair<uint64_t, uint64_t>, std::vector<edm_problem *>> byshape;
        for (edm_problem *p : act) {
            if (p->op->fast && p->op->cfwd.ok)
                byshape[{p->b->id, p->op->cfwd.shape_hash}].push_back(p);
            else
                solo.push_back(p);
        }
        // groups: runs of up to maxg problems of one (basis, structure); large ones gather x
        // from an interleaved copy of the group's vectors (interleave jobs, run before the
        // shared walk), carved from one arena per step
        std::vector<std::pair<edm_problem *const *, int>> groups;
        for (auto &kv : byshape) {
            std::vector<edm_problem *> &v = kv.second;
            for (size_t k = 0; k < v.size();) {
                const int G = (int)std::min<size_t>((size_t)maxg, v.size() - k);
                if (G == 1) solo.push_back(v[k]);
                else groups.emplace_back(v.data() + k, G);
                k += (size_t)G;
            }
        }
        const int64_t imin = batch_interleave_min();
        size_t xi_bytes = 0;
        for (auto &gr : groups)
            if (imin >= 0 && gr.first[0]->D >= imin)
                xi_bytes += align_up((size_t)gr.second * gr.first[0]->D * sizeof(double2));
        bool inter = xi_bytes > 0;
        if (inter) {
            try {
                bt->d_xi.ensure(ctx, xi_bytes);
            } catch (const EdmError &e) {
                if (e.code != EDM_ENOMEM) throw;
                inter = false;   // no room for the copies: gather from the vectors themselves
            }
        }
        bt->xjobs.clear();
        size_t xi_at = 0;
        const edm_op *ops[8];
        const double2 *xs[8];
        double2 *ys[8];
        for (auto &gr : groups) {
            edm_problem *const *v = gr.first;
            const int G = gr.second;
            for (int g = 0; g < G; ++g) {
                ops[g] = v[g]->op;
                xs[g] = v[g]->V[v[g]->j];
                ys[g] = v[g]->V[v[g]->j + 1];
            }
            double2 *xi = nullptr;
            if (inter && v[0]->D >= imin) {
                xi = (double2 *)((unsigned char *)bt->d_xi.p + xi_at);
                xi_at += align_up((size_t)G * v[0]->D * sizeof(double2));
            }
            const bool big = big_groups && v[0]->D >= big_rows;
            std::vector<unsigned char> &dv = big ? bt->gdesc_big[G] : bt->gdesc[G];
            std::vector<int> &bv = big ? bt->gblk_big[G] : bt->gblk[G];
            const size_t at = dv.size();
            dv.resize(at + gsz);
            size_t l = 0;
            if (batch_mv_group_desc(ops, v[0]->b, xs, ys, G, (int64_t)bv.size(), dv.data() + at,
                                    &l, xi, big)) {
                bv.insert(bv.end(), (size_t)batch_mv_blocks(v[0]->D, big ? 1024 : 512),
                          (int)(at / gsz));
                if (big) lds_big[G] = std::max(lds_big[G], l);
                else lds = std::max(lds, l);
                if (xi) {
                    XJob job{G, v[0]->D, {}, xi};
                    for (int g = 0; g < G; ++g) job.x[g] = xs[g];
                    bt->xjobs.push_back(job);
                }
            } else {
                dv.resize(at);
                for (int g = 0; g < G; ++g) solo.push_back(v[g]);
            }
        }
    }
    for (edm_problem *p : solo) {
        size_t l = 0;
        const size_t at = mvdesc.size();
        mvdesc.resize(at + mvsz);
        int64_t blk0 = (int64_t)mvblk.size();
        if (batch_mv_desc(p->op, p->b, p->V[p->j], p->V[p->j + 1], blk0, mvdesc.data() + at,
                          &l)) {
            int pid = (int)(at / mvsz);
            mvblk.insert(mvblk.end(), (size_t)batch_mv_blocks(p->D, 512), pid);
            lds = std::max(lds, l);
        } else {
            mvdesc.resize(at);
            single_mv.push_back(p);
        }
    }
    size_t mv_off = pl.put(mvdesc.data(), mvdesc.size());
    size_t mvblk_off = pl.put(mvblk.data(), mvblk.size() * sizeof(int));
    size_t gdesc_off[8] = {0}, gblk_off[8] = {0}, gbdesc_off[8] = {0}, gbblk_off[8] = {0};
    for (int g = 2; g <= maxg; ++g) {
        gbdesc_off[g] = pl.put(bt->gdesc_big[g].data(), bt->gdesc_big[g].size());
        gbblk_off[g] = pl.put(bt->gblk_big[g].data(), bt->gblk_big[g].size() * sizeof(int));
        gdesc_off[g] = pl.put(bt->gdesc[g].data(), bt->gdesc[g].size());
        gblk_off[g] = pl.put(bt->gblk[g].data(), bt->gblk[g].size() * sizeof(int));
    }
    std::vector<size_t> &vtab_off = bt->vtab_off;
    vtab_off.resize(n);
    for (int q = 0; q < n; ++q)
        vtab_off[q] = pl.put(act[q]->V.data(), (size_t)(act[q]->j + 1) * sizeof(double2 *));
    std::vector<BatchVecDesc> &vd = bt->vd;
    vd.resize(n);
    std::vector<int> &tile2prob = bt->tile2prob, &blk2prob = bt->blk2prob;
    tile2prob.clear();
    blk2prob.clear();
    int maxvec = 0;
    unsigned char *dplan_base = nullptr;   // filled after the arena is sized
    unsigned char *dcoef = nullptr;
    for (int q = 0; q < n; ++q) {
        edm_problem *p = act[q];
        BatchVecDesc &d = vd[q];
        d.D = p->D;
        d.nvec = p->j + 1;
        d.w = p->V[p->j + 1];
        d.partial = p->partial;
        d.ntiles = p->ntiles;
        d.tile0 = (int64_t)tile2prob.size();
        d.blk0 = (int64_t)blk2prob.size();
        tile2prob.insert(tile2prob.end(), (size_t)p->ntiles, q);
        blk2prob.insert(blk2prob.end(), (size_t)rblocks_of(p->D), q);
        maxvec = std::max(maxvec, d.nvec);
    }
    size_t vd_off = pl.reserve(n * sizeof(BatchVecDesc));
    size_t t2p_off = pl.put(tile2prob.data(), tile2prob.size() * sizeof(int));
    size_t b2p_off = pl.put(blk2prob.data(), blk2prob.size() * sizeof(int));
    bt->d_plan.ensure(ctx, pl.h.size());
    dplan_base = (unsigned char *)bt->d_plan.p;
    dcoef = (unsigned char *)bt->d_coef.p;
    for (int q = 0; q < n; ++q) {
        BatchVecDesc &d = vd[q];
        d.V = (const double2 *const *)(dplan_base + vtab_off[q]);
        d.h = (double2 *)(dcoef + coef_off[q]);
        d.h2 = d.h + (act[q]->m + 1);
        d.beta = (double *)(d.h2 + (act[q]->m + 1));
    }
    if (n > 0) {
        memcpy(pl.h.data() + vd_off, vd.data(), n * sizeof(BatchVecDesc));
        bt->h_plan.ensure(st, pl.h.size());
        memcpy(bt->h_plan.p, pl.h.data(), pl.h.size());
        EDM_HIP_CHECK(hipMemcpyAsync(bt->d_plan.p, bt->h_plan.p, pl.h.size(),
                                     hipMemcpyHostToDevice, st));
        // device work of the step
        for (edm_problem *p : single_mv)
            launch_matvec(ctx, p->op, p->b, p->V[p->j], p->V[p->j + 1], ApplyMode::Forward);
        launch_batch_mv(ctx, dplan_base + mv_off, (const int *)(dplan_base + mvblk_off),
                        (int64_t)mvblk.size(), lds);
        for (const XJob &job : bt->xjobs) launch_batch_interleave(ctx, job.G, job.x, job.D, job.xi);
        for (int g = 2; g <= maxg; ++g)
            launch_batch_mv_group(ctx, g, dplan_base + gdesc_off[g],
                                  (const int *)(dplan_base + gblk_off[g]),
                                  (int64_t)bt->gblk[g].size(), lds, 512);
        for (int g = 2; g <= maxg; ++g)
            launch_batch_mv_group(ctx, g, dplan_base + gbdesc_off[g],
                                  (const int *)(dplan_base + gbblk_off[g]),
                                  (int64_t)bt->gblk_big[g].size(), lds_big[g], 1024);
        const BatchVecDesc *dvd = (const BatchVecDesc *)(dplan_base + vd_off);
        bblas_cgs2_norm(ctx, dvd, n, maxvec, (const int *)(dplan_base + t2p_off),
                        (int64_t)tile2prob.size(), (const int *)(dplan_base + b2p_off),
                        (int64_t)blk2prob.size());
        bblas_scale(ctx, dvd, (const int *)(dplan_base + b2p_off), (int64_t)blk2prob.size());
        EDM_HIP_CHECK(hipMemcpyAsync(bt->h_coef.p, bt->d_coef.p, coef_bytes,
                                     hipMemcpyDeviceToHost, st));
        ++bt->steps;
    }

    // while the device works: Rayleigh-Ritz of the parked problems (independent, in
    // parallel), then their restart / final-vector rotations, queued behind the step
    std::vector<std::pair<edm_problem *, int>> rots, restarts;
    std::vector<edm_problem *> finishing;
    std::vector<int> ncols(parked.size());
    std::vector<char> fins(parked.size());
    bt->pool.run((int)parked.size(), [&](int i) {
        bool fin = false;
        ncols[i] = host_cycle_end(parked[i], &fin);
        fins[i] = fin;
    });
    for (size_t i = 0; i < parked.size(); ++i) {
        edm_problem *p = parked[i];
        if (fins[i]) {
            finishing.push_back(p);
            if (ncols[i] > 0) rots.push_back({p, ncols[i]});
        } else {
            rots.push_back({p, ncols[i]});
            restarts.push_back({p, ncols[i]});
        }
    }
    queue_rotations(bt, rots);
    for (auto &r : restarts) host_restart(r.first, r.second);
    EDM_HIP_CHECK(hipStreamSynchronize(st));
    for (edm_problem *p : finishing) finish_problem(p);

    // coefficients of the stepped problems (serial: a breakdown launches device work)
    std::vector<double2> h;
    for (int q = 0; q < n; ++q) {
        edm_problem *p = act[q];
        const double2 *hc = (const double2 *)(bt->h_coef.p + coef_off[q]);
        const double2 *h2c = hc + (p->m + 1);
        double beta = *(const double *)(h2c + (p->m + 1));
        int nv = p->j + 1;
        h.resize(nv);
        for (int i = 0; i < nv; ++i) h[i] = make_double2(hc[i].x + h2c[i].x, hc[i].y + h2c[i].y);
        p->matvecs += 1;
        host_step(ctx, p, h, beta);
    }
    return (int)finishing.size();
}

static edm_problem *find_problem(edm_batch *bt, int64_t id)
{
    auto it = bt->probs.find(id);
    EDM_REQUIRE(it != bt->probs.end(), EDM_EINVAL, "unknown problem id");
    return it->second;
}

// ---------------------------------------------------------------------------- entry helpers
edm_batch *batch_create(edm_ctx *ctx)
{
    edm_batch *bt = new edm_batch();
    bt->ctx = ctx;
    return bt;
}

void batch_destroy(edm_batch *bt)
{
    edm_ctx *ctx = bt->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : bt->probs) release_problem(bt, kv.second);
    bt->probs.clear();
    bt->slabs.release_all(ctx);
    bt->d_plan.release(ctx);
    bt->d_coef.release(ctx);
    bt->d_xi.release(ctx);
    bt->d_rot.release(ctx);
    delete bt;
}

int batch_step(edm_batch *bt, int max_steps, int *n_running, int *n_finished)
{
    EDM_REQUIRE(max_steps >= 1, EDM_EINVAL, "max_steps must be at least 1");
    int fin = 0;
    for (int s = 0; s < max_steps; ++s) {
        fin += batch_step_once(bt);
        if (fin > 0) break;
        bool any = false;
        for (auto &kv : bt->probs) any |= kv.second->state == kRunning;
        if (!any) break;
    }
    int run = 0;
    for (auto &kv : bt->probs) run += kv.second->state == kRunning;
    if (n_running) *n_running = run;
    if (n_finished) *n_finished = fin;
    return EDM_OK;
}

int batch_finished(edm_batch *bt, int64_t *ids, int cap, int *n)
{
    int c = 0;
    for (auto &kv : bt->probs) {
        if (kv.second->state == kRunning) continue;
        if (ids && c < cap) ids[c] = kv.first;
        ++c;
    }
    *n = c;
    return EDM_OK;
}

int batch_status(edm_batch *bt, int64_t id, int *state, edm_solve_info *info)
{
    edm_problem *p = find_problem(bt, id);
    if (state) *state = p->state;
    if (info) {
        info->iterations = std::min(p->iters, p->maxiter);
        info->matvecs = p->matvecs;
        info->converged = p->nconv;
        info->krylovdim = p->m;
        double maxres = 0.0;
        if (p->state != kRunning)
            for (int i = 0; i < p->nev; ++i) maxres = std::max(maxres, p->res[i]);
        info->max_residual = maxres;
        info->seconds = p->state == kRunning
            ? std::chrono::duration<double>(std::chrono::steady_clock::now() - p->t0).count()
            : p->seconds;
    }
    return EDM_OK;
}

int batch_result(edm_batch *bt, int64_t id, double *vals, double *vecs, edm_solve_info *info)
{
    edm_ctx *ctx = bt->ctx;
    edm_problem *p = find_problem(bt, id);
    EDM_REQUIRE(p->state != kRunning, EDM_EINVAL, "problem is still running");
    EDM_REQUIRE(!vecs || p->want_vectors, EDM_EINVAL,
                "eigenvectors were not requested when the problem was added");
    batch_status(bt, id, nullptr, info);
    if (vals)
        for (int i = 0; i < p->nev; ++i) vals[i] = p->vals[i];
    if (vecs) {
        for (int i = 0; i < p->nev; ++i)
            EDM_HIP_CHECK(hipMemcpyAsync(vecs + 2 * (size_t)i * p->D, p->tmp[i],
                                         p->D * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    int rc = p->state == kDone ? EDM_OK : EDM_ENOCONV;
    bt->probs.erase(id);
    release_problem(bt, p);
    return rc;
}

int batch_remove(edm_batch *bt, int64_t id)
{
    edm_ctx *ctx = bt->ctx;
    edm_problem *p = find_problem(bt, id);
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    bt->probs.erase(id);
    release_problem(bt, p);
    return EDM_OK;
}

int batch_count(edm_batch *bt, int *n_total, int *n_running)
{
    int run = 0;
    for (auto &kv : bt->probs) run += kv.second->state == kRunning;
    if (n_total) *n_total = (int)bt->probs.size();
    if (n_running) *n_running = run;
    return EDM_OK;
}

int batch_set_weight(edm_batch *bt, int64_t id, double weight)
{
    EDM_REQUIRE(weight > 0 && std::isfinite(weight), EDM_EINVAL, "weight must be positive");
    edm_problem *p = find_problem(bt, id);
    p->weight = weight;
    return EDM_OK;
}

int batch_set_quantum(edm_batch *bt, int64_t rows)
{
    EDM_REQUIRE(rows >= 0, EDM_EINVAL, "quantum must be >= 0");
    bt->quantum = rows;
    for (auto &kv : bt->probs) kv.second->credit = 0;
    return EDM_OK;
}

uint64_t batch_problem_bytes(int64_t D, int nev, int krylovdim)
{
    int m = krylov_dim(D, nev, krylovdim);
    return (uint64_t)align_up(problem_slab_bytes(D, std::max(m, 1)));
}

edm_ctx *batch_ctx(edm_batch *bt) { return bt->ctx; }
