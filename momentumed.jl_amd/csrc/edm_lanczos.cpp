// edm_lanczos.cpp — thick-restart Lanczos for the lowest eigenpairs of a Hermitian map.
//
// Stands in for KrylovKit.eigsolve(A, x0, howmany, :SR; ishermitian=true, krylovdim, tol,
// maxiter) as called by krylov_map_solve (MomentumED/src/method/linear_map.jl:179-191,
// MomentumED/ext/MomentumEDCUDAExt.jl:159-179). Algorithm: the thick-restart Lanczos method
// of Wu & Simon (SIAM J. Matrix Anal. Appl. 22, 2000) with full re-orthogonalisation by
// classical Gram-Schmidt applied twice ("twice is enough", Giraud, Langou & Rozloznik 2005).
// Every vector lives on the device; per Lanczos step the host reads back only the projection
// coefficients and the norm (a few hundred bytes). The projected matrix (tridiagonal plus an
// arrow after each restart) is real symmetric and is diagonalised on the host (sym_eigh:
// Householder tridiagonalisation + implicit QL).
//
// Restart: keep the `keep` lowest Ritz vectors, keep = (3*krylovdim + 2*converged) / 5 (at
// least the number of wanted pairs), rotate them into place with one tall-skinny product,
// and continue from the normalised residual. Convergence of Ritz pair i: |beta_m * Y[m-1,i]|
// <= tol (KrylovKit's absolute residual criterion).
#include "edm_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>

namespace {

// Implicit QL iteration with Wilkinson shifts on the real symmetric tridiagonal matrix
// (d, e) (e[i] couples i and i+1), each plane rotation applied to the columns of the
// accumulated basis Qt (column c at Qt[c*n .. c*n+n), real or complex entries) (Bowdler,
// Martin, Reinsch & Wilkinson, Numer. Math. 11, 1968). On return d holds the eigenvalues
// (unsorted) and column c of Qt the eigenvector of d[c].
template <class T>
void tridiag_ql(int n, std::vector<double> &d, std::vector<double> &e, std::vector<T> &Qt)
{
    auto hyp = [](double x, double y) { return std::sqrt(x * x + y * y); };
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        for (int iter = 0; iter < 200; ++iter) {
            int m = l;
            for (; m + 1 < n; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            // Wilkinson shift from the leading 2x2 block of the unreduced part
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hyp(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                double f = s * e[i], bb = c * e[i];
                r = hyp(f, g);
                e[i + 1] = r;
                if (r == 0.0) {   // exact split: undo the pending shift and restart
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
                T *ci = &Qt[(size_t)i * n], *ci1 = &Qt[(size_t)(i + 1) * n];
                for (int k2 = 0; k2 < n; ++k2) {   // columns i, i+1 of the basis
                    T zi = ci[k2], zi1 = ci1[k2];
                    ci1[k2] = s * zi + c * zi1;
                    ci[k2] = c * zi - s * zi1;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
}

// exponent e with 2^e >= max |x|: scaling by 2^-e is exact and keeps sqrt(f^2 + g^2)
// away from overflow and underflow
int scale_exponent(double amax)
{
    int ex = 0;
    if (amax > 0.0) (void)std::frexp(amax, &ex);
    return ex;
}

}  // namespace

// Eigen-decomposition of a real symmetric n x n matrix A (row-major, destroyed): w[i]
// ascending and Z[:, i] (column-major, ldz = n) the matching orthonormal eigenvectors.
// Householder reduction to tridiagonal form with the reflections accumulated (Golub & Van
// Loan, Matrix Computations, 4th ed., Alg. 8.3.1), then the implicit QL iteration
// (tridiag_ql). O(n^3) with small constants: 0.03 ms at n = 30, the Rayleigh-Ritz step of
// every Lanczos restart.
void sym_eigh(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &Z)
{
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    double amax = 0.0;
    for (double x : A) amax = std::max(amax, std::fabs(x));
    const int ex = scale_exponent(amax);
    amax = std::ldexp(1.0, ex);
    const double inv = std::ldexp(1.0, -ex);
    for (double &x : A) x *= inv;
    // the accumulated orthogonal basis, stored by columns: qt(c, r) = Q[r][c]
    std::vector<double> Qt((size_t)n * n, 0.0);
    auto qt = [&](int c, int r) -> double & { return Qt[(size_t)c * n + r]; };
    for (int i = 0; i < n; ++i) qt(i, i) = 1.0;
    std::vector<double> v(n), pv(n), sr(n);
    for (int k = 0; k + 2 < n; ++k) {
        // reflect x = A[k+1:, k] onto -sign(x0) |x| e1
        double xn = 0.0;
        for (int i = k + 1; i < n; ++i) xn += a(i, k) * a(i, k);
        xn = std::sqrt(xn);
        double tail = 0.0;
        for (int i = k + 2; i < n; ++i) tail += a(i, k) * a(i, k);
        if (tail == 0.0) continue;   // already tridiagonal in this column
        double x0 = a(k + 1, k);
        double alpha = x0 >= 0.0 ? -xn : xn;
        for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
        v[k + 1] -= alpha;
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        double beta = 2.0 / vv;
        // A22 <- H A22 H with H = I - beta v v^T:  p = beta A22 v, K = beta/2 v.p,
        // u = p - K v, A22 -= v u^T + u v^T
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j];
            pv[i] = beta * s;
            vp += v[i] * pv[i];
        }
        double K = 0.5 * beta * vp;
        for (int i = k + 1; i < n; ++i) pv[i] -= K * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) a(i, j) -= v[i] * pv[j] + pv[i] * v[j];
        a(k + 1, k) = alpha;
        a(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
        // Q <- Q H: Q[r][:] -= beta (Q[r][:] . v) v^T
        std::fill(sr.begin(), sr.end(), 0.0);
        for (int j = k + 1; j < n; ++j) {
            const double vj = v[j];
            const double *col = &Qt[(size_t)j * n];
            for (int r = 0; r < n; ++r) sr[r] += col[r] * vj;
        }
        for (int j = k + 1; j < n; ++j) {
            const double bvj = beta * v[j];
            double *col = &Qt[(size_t)j * n];
            for (int r = 0; r < n; ++r) col[r] -= sr[r] * bvj;
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    tridiag_ql(n, d, e, Qt);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    Z.assign((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        w[c] = d[idx[c]] * amax;
        std::copy(Qt.begin() + (size_t)idx[c] * n, Qt.begin() + (size_t)idx[c] * n + n,
                  Z.begin() + (size_t)c * n);
    }
}

// Eigen-decomposition of a complex Hermitian n x n matrix A (row-major, destroyed): w[i]
// ascending and Z[:, i] (column-major) orthonormal eigenvectors. Householder reduction with
// complex reflectors H = I - beta v v^H (beta real) to a Hermitian tridiagonal matrix, whose
// off-diagonal phases are then moved into the basis by a diagonal unitary (Golub & Van Loan,
// 4th ed., Sec. 8.3.1 and Problem 8.3.7), leaving a real symmetric tridiagonal matrix for
// the implicit QL iteration.
void herm_eigh(int n, std::vector<std::complex<double>> &A, std::vector<double> &w,
               std::vector<std::complex<double>> &Z)
{
    using cd = std::complex<double>;
    auto a = [&](int i, int j) -> cd & { return A[(size_t)i * n + j]; };
    double amax = 0.0;
    for (const cd &x : A) amax = std::max(amax, std::max(std::fabs(x.real()), std::fabs(x.imag())));
    const int ex = scale_exponent(amax);
    amax = std::ldexp(1.0, ex);
    const double inv = std::ldexp(1.0, -ex);
    for (cd &x : A) x *= inv;
    for (int i = 0; i < n; ++i) a(i, i) = cd(a(i, i).real(), 0.0);
    std::vector<cd> Qt((size_t)n * n, cd(0.0, 0.0));
    auto qt = [&](int c, int r) -> cd & { return Qt[(size_t)c * n + r]; };
    for (int i = 0; i < n; ++i) qt(i, i) = 1.0;
    std::vector<cd> v(n), pv(n), sr(n);
    for (int k = 0; k + 2 < n; ++k) {
        // reflect x = A[k+1:, k] onto alpha e1, alpha = -(x0 / |x0|) |x|
        double xn = 0.0, tail = 0.0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(a(i, k));
        for (int i = k + 2; i < n; ++i) tail += std::norm(a(i, k));
        xn = std::sqrt(xn);
        if (tail == 0.0) continue;
        const cd x0 = a(k + 1, k);
        const double ax0 = std::abs(x0);
        const cd ph = ax0 > 0.0 ? x0 / ax0 : cd(1.0, 0.0);
        const cd alpha = -ph * xn;
        for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
        v[k + 1] -= alpha;
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += std::norm(v[i]);
        const double beta = 2.0 / vv;
        // A22 <- H A22 H:  p = beta A22 v, K = beta/2 v^H p (real), u = p - K v,
        // A22 -= v u^H + u v^H
        cd vp = 0.0;
        for (int i = k + 1; i < n; ++i) {
            cd s = 0.0;
            for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j];
            pv[i] = beta * s;
            vp += std::conj(v[i]) * pv[i];
        }
        const double K = 0.5 * beta * vp.real();
        for (int i = k + 1; i < n; ++i) pv[i] -= K * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j)
                a(i, j) -= v[i] * std::conj(pv[j]) + pv[i] * std::conj(v[j]);
        for (int i = k + 1; i < n; ++i) a(i, i) = cd(a(i, i).real(), 0.0);
        a(k + 1, k) = alpha;
        a(k, k + 1) = std::conj(alpha);
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
        // Q <- Q H: Q[r][j] -= beta (sum_l Q[r][l] v_l) conj(v_j)
        std::fill(sr.begin(), sr.end(), cd(0.0, 0.0));
        for (int j = k + 1; j < n; ++j) {
            const cd vj = v[j];
            const cd *col = &Qt[(size_t)j * n];
            for (int r = 0; r < n; ++r) sr[r] += col[r] * vj;
        }
        for (int j = k + 1; j < n; ++j) {
            const cd bvj = beta * std::conj(v[j]);
            cd *col = &Qt[(size_t)j * n];
            for (int r = 0; r < n; ++r) col[r] -= sr[r] * bvj;
        }
    }
    // T = P T' P^H with P = diag(ph), T' real: ph[0] = 1, ph[i+1] = ph[i] e_i / |e_i|;
    // the eigenvectors of A are Q P z
    std::vector<double> d(n), e(n, 0.0);
    cd phase = 1.0;
    for (int i = 0; i < n; ++i) {
        d[i] = a(i, i).real();
        if (i > 0) {
            cd* col = &Qt[(size_t)i * n];
            for (int r = 0; r < n; ++r) col[r] *= phase;
        }
        if (i + 1 < n) {
            const cd ei = a(i + 1, i);
            const double ae = std::abs(ei);
            e[i] = ae;
            if (ae > 0.0) phase *= ei / ae;
        }
    }
    tridiag_ql(n, d, e, Qt);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    Z.assign((size_t)n * n, cd(0.0, 0.0));
    for (int c = 0; c < n; ++c) {
        w[c] = d[idx[c]] * amax;
        std::copy(Qt.begin() + (size_t)idx[c] * n, Qt.begin() + (size_t)idx[c] * n + n,
                  Z.begin() + (size_t)c * n);
    }
}

namespace {

struct Workspace {
    edm_ctx *ctx;
    int64_t D;
    int m;
    std::vector<double2 *> V, tmp;
    double2 **d_vptrs = nullptr, **d_tmpptrs = nullptr;
    double2 *d_h = nullptr, *d_h2 = nullptr;
    double *d_beta = nullptr, *d_Y = nullptr;
    // two pinned landing slots for a step's h, h2 and beta, each with the event recorded
    // after its copy, so the host can wait for one step while the next is already queued
    double2 *h_pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t vbytes = 0;

    size_t coef_count() const { return 2 * (size_t)(m + 1) + 1; }

    Workspace(edm_ctx *c, int64_t D_, int m_, int ntmp) : ctx(c), D(D_), m(m_)
    {
        vbytes = (size_t)std::max<int64_t>(D, 1) * sizeof(double2);
        try {
            for (int i = 0; i <= m; ++i) V.push_back((double2 *)ctx->alloc(vbytes));
            for (int i = 0; i < ntmp; ++i) tmp.push_back((double2 *)ctx->alloc(vbytes));
            d_vptrs = (double2 **)ctx->alloc((m + 1) * sizeof(double2 *));
            d_tmpptrs = (double2 **)ctx->alloc(std::max(ntmp, 1) * sizeof(double2 *));
            d_h = (double2 *)ctx->alloc((2 * (m + 1) + 1) * sizeof(double2));
            d_h2 = d_h + (m + 1);
            d_beta = (double *)(d_h2 + (m + 1));
            d_Y = (double *)ctx->alloc((size_t)m * m * sizeof(double));
            for (int s = 0; s < 2; ++s) {
                EDM_HIP_CHECK(hipHostMalloc((void **)&h_pin[s], coef_count() * sizeof(double2),
                                            hipHostMallocDefault));
                EDM_HIP_CHECK(hipEventCreateWithFlags(&ev[s], hipEventDisableTiming));
            }
        } catch (...) {
            release();
            throw;
        }
        upload_ptrs();
    }
    void upload_ptrs()
    {
        EDM_HIP_CHECK(hipMemcpyAsync(d_vptrs, V.data(), V.size() * sizeof(double2 *),
                                     hipMemcpyHostToDevice, ctx->stream));
        if (!tmp.empty())
            EDM_HIP_CHECK(hipMemcpyAsync(d_tmpptrs, tmp.data(), tmp.size() * sizeof(double2 *),
                                         hipMemcpyHostToDevice, ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    void release()
    {
        // A step queued ahead may still be running when an error unwinds: it writes V[j+2],
        // d_h and a pinned slot. Wait for the stream before any of them is freed.
        (void)hipStreamSynchronize(ctx->stream);
        for (auto p : V) ctx->free(p, vbytes);
        for (auto p : tmp) ctx->free(p, vbytes);
        V.clear();
        tmp.clear();
        ctx->free(d_vptrs, (m + 1) * sizeof(double2 *));
        ctx->free(d_tmpptrs, sizeof(double2 *));
        ctx->free(d_h, (2 * (m + 1) + 1) * sizeof(double2));
        ctx->free(d_Y, (size_t)m * m * sizeof(double));
        for (int s = 0; s < 2; ++s) {
            if (h_pin[s]) (void)hipHostFree(h_pin[s]);
            if (ev[s]) (void)hipEventDestroy(ev[s]);
            h_pin[s] = nullptr;
            ev[s] = nullptr;
        }
        d_vptrs = d_tmpptrs = nullptr;
        d_h = d_h2 = nullptr;
        d_beta = d_Y = nullptr;
    }
    ~Workspace() { release(); }
};

// Orthogonalise V[target] against V[0..n) twice; returns the first-pass coefficients + second
// pass corrections summed on the host (h_out[n]).
void cgs2(Workspace &ws, int n, double2 *target, std::vector<double2> &h_out)
{
    edm_ctx *ctx = ws.ctx;
    h_out.assign(n, make_double2(0.0, 0.0));
    if (n == 0) return;
    std::vector<double2> h2(n);
    blas_cgs2(ctx, ws.d_vptrs, n, target, ws.D, ws.d_h, ws.d_h2);
    EDM_HIP_CHECK(hipMemcpyAsync(h_out.data(), ws.d_h, n * sizeof(double2), hipMemcpyDeviceToHost,
                                 ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(h2.data(), ws.d_h2, n * sizeof(double2), hipMemcpyDeviceToHost,
                                 ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
        h_out[i].x += h2[i].x;
        h_out[i].y += h2[i].y;
    }
}

// Queue CGS2 of w against V[0..n), then beta = ||w|| and w /= beta, then one copy of h, h2
// and beta into pinned slot `slot`, whose event marks the copy's completion.
void queue_cgs2_norm_scale(Workspace &ws, int n, double2 *w, int slot)
{
    edm_ctx *ctx = ws.ctx;
    blas_cgs2(ctx, ws.d_vptrs, n, w, ws.D, ws.d_h, ws.d_h2);
    blas_norm(ctx, w, ws.D, ws.d_beta);
    blas_scale_inv(ctx, w, w, ws.d_beta, ws.D);
    // h, h2 and beta are adjacent on the device (d_h2 = d_h + m + 1, d_beta follows). Work
    // queued after this copy may overwrite them: the stream runs the copy first.
    EDM_HIP_CHECK(hipMemcpyAsync(ws.h_pin[slot], ws.d_h, ws.coef_count() * sizeof(double2),
                                 hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipEventRecord(ws.ev[slot], ctx->stream));
}

// Wait for slot `slot`'s copy only (work queued behind it keeps running); h_out = first-pass
// + second-pass coefficients of the n basis vectors.
void collect_cgs2_norm_scale(Workspace &ws, int n, int slot, std::vector<double2> &h_out,
                             double *beta)
{
    EDM_HIP_CHECK(hipEventSynchronize(ws.ev[slot]));
    const double2 *buf = ws.h_pin[slot];
    const double2 *h2 = buf + (ws.m + 1);
    h_out.resize(n);
    for (int i = 0; i < n; ++i) h_out[i] = make_double2(buf[i].x + h2[i].x, buf[i].y + h2[i].y);
    *beta = *(const double *)(buf + 2 * (ws.m + 1));
}

double norm_host(Workspace &ws, const double2 *v)
{
    blas_norm(ws.ctx, v, ws.D, ws.d_beta);
    double b = 0.0;
    EDM_HIP_CHECK(hipMemcpyAsync(&b, ws.d_beta, sizeof(double), hipMemcpyDeviceToHost,
                                 ws.ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ws.ctx->stream));
    return b;
}

}  // namespace

int run_eigsolve(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int nev, int krylovdim,
                 double tol, int maxiter, uint64_t seed, const double *vec0, double *vals,
                 double *vecs, edm_solve_info *info)
{
    return run_eigsolve_mv(
        ctx, b->D,
        [&](const double2 *x, double2 *y) { launch_matvec(ctx, op, b, x, y, ApplyMode::Forward); },
        nev, krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
}

int run_eigsolve_mv(edm_ctx *ctx, int64_t D, const std::function<void(const double2 *, double2 *)> &mv,
                    int nev, int krylovdim, double tol, int maxiter, uint64_t seed,
                    const double *vec0, double *vals, double *vecs, edm_solve_info *info)
{
    auto t0 = std::chrono::steady_clock::now();
    EDM_REQUIRE(D > 0, EDM_EINVAL, "empty Hilbert subspace");
    EDM_REQUIRE(nev >= 1, EDM_EINVAL, "N should be at least 1.");
    EDM_REQUIRE(nev <= D, EDM_EINVAL, "more eigenpairs requested than the subspace dimension");
    EDM_REQUIRE(tol > 0, EDM_EINVAL, "tol must be positive");
    EDM_REQUIRE(maxiter >= 1, EDM_EINVAL, "maxiter must be at least 1");
    int m = std::max(krylovdim, nev + 2);
    if ((int64_t)m > D) m = (int)D;
    EDM_REQUIRE(m <= 255, EDM_EINVAL, "krylovdim must be at most 255");
    int ntmp = m;
    Workspace ws(ctx, D, m, ntmp);
    // size the reduction scratch once for the largest multi-dot
    {
        int64_t nb = (D + 1023) / 1024;
        ctx->partials((size_t)(nb + 1) * (m + 1) * sizeof(double2));
    }
    hipStream_t st = ctx->stream;
    int matvecs = 0;
    uint64_t rng_stream = 0;

    // start vector
    if (vec0) {
        EDM_HIP_CHECK(hipMemcpyAsync(ws.V[0], vec0, D * sizeof(double2), hipMemcpyHostToDevice, st));
    } else {
        blas_fill_random(ctx, ws.V[0], D, seed, 0);
    }
    double b0 = norm_host(ws, ws.V[0]);
    EDM_REQUIRE(b0 > 0 && std::isfinite(b0), EDM_EINVAL, "start vector is zero or not finite");
    blas_scale_inv(ctx, ws.V[0], ws.V[0], ws.d_beta, D);

    std::vector<double> T((size_t)m * m, 0.0);
    auto t = [&](int i, int j) -> double & { return T[(size_t)i * m + j]; };
    std::vector<double> theta, Y;
    std::vector<double2> h;
    int k = 0;            // kept Ritz vectors at the start of a cycle
    int iters = 0;
    int nconv = 0;
    double beta_last = 0.0;
    double hnorm = 0.0;   // running estimate of ||H|| for the breakdown threshold
    std::vector<double> res(m, 0.0);

    for (iters = 1; iters <= maxiter; ++iters) {
        // A step queues the matvec, both Gram-Schmidt passes, the norm, the scaling and the
        // copy of h, h2 and beta. Its successor only needs the scaled vector, which is on the
        // device, so it is queued before the host waits for this step's coefficients: the
        // device does not idle during the host's round trip. When a step breaks down, the
        // step queued behind it is discarded (its vector and coefficients are written again).
        bool queued = false;   // step j is already on the stream
        int slot = 0;          // pinned slot of step j's coefficients
        for (int j = k; j < m; ++j) {
            if (!queued) {
                mv(ws.V[j], ws.V[j + 1]);
                queue_cgs2_norm_scale(ws, j + 1, ws.V[j + 1], slot);
            }
            queued = j + 1 < m;
            if (queued) {
                mv(ws.V[j + 1], ws.V[j + 2]);
                queue_cgs2_norm_scale(ws, j + 2, ws.V[j + 2], slot ^ 1);
            }
            ++matvecs;
            double beta = 0.0;
            collect_cgs2_norm_scale(ws, j + 1, slot, h, &beta);
            slot ^= 1;
            t(j, j) = h[j].x;
            hnorm = std::max(hnorm, std::fabs(h[j].x) + beta + (j > 0 ? std::fabs(t(j, j - 1)) : 0.0));
            bool breakdown = !(beta > 1e-14 * std::max(hnorm, 1e-300));
            if (breakdown) {
                // invariant subspace: continue from a fresh random direction orthogonal to
                // the current basis (coupling zero); the step queued behind this one ran on
                // the unusable vector and is run again
                EDM_HIP_CHECK(hipStreamSynchronize(st));
                queued = false;
                beta = 0.0;
                for (int attempt = 0; attempt < 4; ++attempt) {
                    blas_fill_random(ctx, ws.V[j + 1], D, seed ^ (0xA5A5A5A5ULL + ++rng_stream), 0);
                    std::vector<double2> hh;
                    cgs2(ws, j + 1, ws.V[j + 1], hh);
                    double nb = norm_host(ws, ws.V[j + 1]);
                    if (nb > 1e-8) break;
                }
                blas_scale_inv(ctx, ws.V[j + 1], ws.V[j + 1], ws.d_beta, D);
            }
            if (j + 1 < m) {
                t(j, j + 1) = beta;
                t(j + 1, j) = beta;
            }
            beta_last = beta;
        }
        // Rayleigh-Ritz on the projected matrix
        std::vector<double> A = T;
        sym_eigh(m, A, theta, Y);
        nconv = 0;
        for (int i = 0; i < m; ++i) res[i] = std::fabs(beta_last * Y[(size_t)i * m + (m - 1)]);
        while (nconv < nev && res[nconv] <= tol) ++nconv;
        if (nconv >= nev || iters == maxiter || (int64_t)m == D) break;
        // thick restart
        int keep = (3 * m + 2 * nconv) / 5;
        keep = std::max(keep, nev);
        keep = std::min(keep, m - 1);
        EDM_HIP_CHECK(hipMemcpyAsync(ws.d_Y, Y.data(), (size_t)m * keep * sizeof(double),
                                     hipMemcpyHostToDevice, st));
        blas_rotate(ctx, (const double2 *const *)ws.d_vptrs, m, ws.d_Y, m, keep,
                    (double2 *const *)ws.d_tmpptrs, D);
        EDM_HIP_CHECK(hipStreamSynchronize(st));
        for (int c = 0; c < keep; ++c) std::swap(ws.V[c], ws.tmp[c]);
        std::swap(ws.V[keep], ws.V[m]);
        ws.upload_ptrs();
        std::fill(T.begin(), T.end(), 0.0);
        for (int c = 0; c < keep; ++c) {
            t(c, c) = theta[c];
            double cpl = beta_last * Y[(size_t)c * m + (m - 1)];
            t(c, keep) = cpl;
            t(keep, c) = cpl;
        }
        k = keep;
    }
    if (iters > maxiter) iters = maxiter;

    for (int i = 0; i < nev; ++i) vals[i] = theta[i];
    double maxres = 0.0;
    for (int i = 0; i < nev; ++i) maxres = std::max(maxres, res[i]);
    if (vecs) {
        EDM_HIP_CHECK(hipMemcpyAsync(ws.d_Y, Y.data(), (size_t)m * nev * sizeof(double),
                                     hipMemcpyHostToDevice, st));
        blas_rotate(ctx, (const double2 *const *)ws.d_vptrs, m, ws.d_Y, m, nev,
                    (double2 *const *)ws.d_tmpptrs, D);
        for (int i = 0; i < nev; ++i)
            EDM_HIP_CHECK(hipMemcpyAsync(vecs + 2 * (size_t)i * D, ws.tmp[i], D * sizeof(double2),
                                         hipMemcpyDeviceToHost, st));
        EDM_HIP_CHECK(hipStreamSynchronize(st));
    }
    if (info) {
        info->iterations = iters;
        info->matvecs = matvecs;
        info->converged = nconv;
        info->krylovdim = m;
        info->max_residual = maxres;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return nconv >= nev ? EDM_OK : EDM_ENOCONV;
}

// ---------------------------------------------------------------- block Lanczos
// Thick-restart block Lanczos with block size bs (Golub & Underwood 1977; thick restart as in
// Wu & Simon 2000, with the restart coupling a bs x keep block instead of a row). A step
// applies H to the current block of bs basis vectors at once (one launch of the block matvec:
// candidate generation, signs and partner lookups are shared by the bs vectors) and
// orthonormalises the bs results column by column with CGS2 against every earlier basis
// vector, so the projected matrix T = V^H H V is block tridiagonal plus a block arrow after a
// restart, and complex Hermitian (herm_eigh). Residual of Ritz pair i: ||Bn y_i[last block]||
// with Bn the triangular coupling of the next block.
namespace {

struct BlockWorkspace {
    edm_ctx *ctx;
    int64_t D;
    int M, bs, ld;                // basis capacity (multiple of bs), block size, stride of d_c
    std::vector<double2 *> V, tmp;
    double2 **d_vptrs = nullptr, **d_tmpptrs = nullptr;
    double2 *d_c = nullptr;       // per column slot: h (M+bs) | h2 (M+bs) | beta
    double2 *d_Y = nullptr;
    size_t vbytes = 0;

    BlockWorkspace(edm_ctx *c, int64_t D_, int M_, int bs_) : ctx(c), D(D_), M(M_), bs(bs_)
    {
        ld = 2 * (M + bs) + 1;
        vbytes = (size_t)std::max<int64_t>(D, 1) * sizeof(double2);
        try {
            for (int i = 0; i < M + bs; ++i) V.push_back((double2 *)ctx->alloc(vbytes));
            for (int i = 0; i < M; ++i) tmp.push_back((double2 *)ctx->alloc(vbytes));
            d_vptrs = (double2 **)ctx->alloc((M + bs) * sizeof(double2 *));
            d_tmpptrs = (double2 **)ctx->alloc(M * sizeof(double2 *));
            d_c = (double2 *)ctx->alloc((size_t)bs * ld * sizeof(double2));
            d_Y = (double2 *)ctx->alloc((size_t)M * M * sizeof(double2));
        } catch (...) {
            release();
            throw;
        }
        upload_ptrs();
    }
    void upload_ptrs()
    {
        EDM_HIP_CHECK(hipMemcpyAsync(d_vptrs, V.data(), V.size() * sizeof(double2 *),
                                     hipMemcpyHostToDevice, ctx->stream));
        EDM_HIP_CHECK(hipMemcpyAsync(d_tmpptrs, tmp.data(), tmp.size() * sizeof(double2 *),
                                     hipMemcpyHostToDevice, ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    double2 *h(int c) { return d_c + (size_t)c * ld; }
    double2 *h2(int c) { return d_c + (size_t)c * ld + (M + bs); }
    double *beta(int c) { return (double *)(d_c + (size_t)c * ld + 2 * (M + bs)); }
    void release()
    {
        // an error may unwind while a block step's kernels are still queued: let the stream
        // drain before their buffers go back to the allocator
        (void)hipStreamSynchronize(ctx->stream);
        for (auto p : V) ctx->free(p, vbytes);
        for (auto p : tmp) ctx->free(p, vbytes);
        V.clear();
        tmp.clear();
        ctx->free(d_vptrs, (M + bs) * sizeof(double2 *));
        ctx->free(d_tmpptrs, M * sizeof(double2 *));
        ctx->free(d_c, (size_t)bs * ld * sizeof(double2));
        ctx->free(d_Y, (size_t)M * M * sizeof(double2));
        d_vptrs = d_tmpptrs = nullptr;
        d_c = d_Y = nullptr;
    }
    ~BlockWorkspace() { release(); }
};

// queue CGS2 of V[n] against V[0..n), its norm and its scaling into column slot c
void queue_column(BlockWorkspace &ws, int c, int n)
{
    edm_ctx *ctx = ws.ctx;
    double2 *w = ws.V[n];
    blas_cgs2(ctx, ws.d_vptrs, n, w, ws.D, ws.h(c), ws.h2(c));
    blas_norm(ctx, w, ws.D, ws.beta(c));
}

// read back slots 0..nc: hsum[c][i] = h + h2 for i < n0 + c, beta[c]
void read_columns(BlockWorkspace &ws, int nc, int n0, std::vector<std::vector<double2>> &hsum,
                  std::vector<double> &beta)
{
    std::vector<double2> buf((size_t)nc * ws.ld);
    EDM_HIP_CHECK(hipMemcpyAsync(buf.data(), ws.d_c, buf.size() * sizeof(double2),
                                 hipMemcpyDeviceToHost, ws.ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ws.ctx->stream));
    hsum.assign(nc, {});
    beta.assign(nc, 0.0);
    for (int c = 0; c < nc; ++c) {
        const double2 *h = buf.data() + (size_t)c * ws.ld, *h2 = h + (ws.M + ws.bs);
        int n = n0 + c;
        hsum[c].resize(n);
        for (int i = 0; i < n; ++i) hsum[c][i] = make_double2(h[i].x + h2[i].x, h[i].y + h2[i].y);
        beta[c] = *(const double *)(h2 + (ws.M + ws.bs));
    }
}

}  // namespace

int run_eigsolve_block(edm_ctx *ctx, int64_t D, int bs, const BlockMatvec &mv, int nev,
                       int krylovdim, double tol, int maxiter, uint64_t seed, const double *vec0,
                       double *vals, double *vecs, edm_solve_info *info)
{
    using cd = std::complex<double>;
    auto t0 = std::chrono::steady_clock::now();
    EDM_REQUIRE(D > 0, EDM_EINVAL, "empty Hilbert subspace");
    EDM_REQUIRE(nev >= 1, EDM_EINVAL, "N should be at least 1.");
    EDM_REQUIRE(nev <= D, EDM_EINVAL, "more eigenpairs requested than the subspace dimension");
    EDM_REQUIRE(tol > 0, EDM_EINVAL, "tol must be positive");
    EDM_REQUIRE(maxiter >= 1, EDM_EINVAL, "maxiter must be at least 1");
    EDM_REQUIRE(bs >= 1 && bs <= 8, EDM_EINVAL, "block size must be in 1..8");
    int M = std::max(krylovdim, nev + 2 * bs);
    M = (M + bs - 1) / bs * bs;
    if (bs == 1 || (int64_t)(M + bs) > D) {
        // one vector at a time (also when the space is too small for M + bs basis vectors)
        return run_eigsolve_mv(
            ctx, D,
            [&](const double2 *x, double2 *y) {
                const double2 *xs[1] = {x};
                double2 *ys[1] = {y};
                mv(1, xs, ys);
            },
            nev, krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
    }
    EDM_REQUIRE(M + bs <= 256, EDM_EINVAL, "krylovdim + block size must be at most 256");
    BlockWorkspace ws(ctx, D, M, bs);
    {
        int64_t nb = (D + 1023) / 1024;
        ctx->partials((size_t)(nb + 1) * (M + bs) * sizeof(double2));
    }
    hipStream_t st = ctx->stream;
    int matvecs = 0;
    uint64_t rng_stream = 0;
    double hnorm = 0.0;

    // orthonormalise V[n] against V[0..n) with a host check after every pass: a column that
    // lost more than 8 digits to cancellation gets a further CGS2, one that vanishes against
    // ||H|| is replaced by a random direction (coupling zero). hsum gets the coefficients.
    auto careful_column = [&](int n, std::vector<double2> &hsum) -> double {
        hsum.assign(n, make_double2(0.0, 0.0));
        bool replaced = false;
        for (int pass = 0; pass < 4; ++pass) {
            queue_column(ws, 0, n);
            std::vector<std::vector<double2>> hh;
            std::vector<double> bb;
            read_columns(ws, 1, n, hh, bb);
            double hn = 0.0;
            for (int i = 0; i < n; ++i) {
                if (!replaced) {
                    hsum[i].x += hh[0][i].x;
                    hsum[i].y += hh[0][i].y;
                }
                hn += hh[0][i].x * hh[0][i].x + hh[0][i].y * hh[0][i].y;
            }
            double beta = bb[0];
            double wn = std::sqrt(hn + beta * beta);
            if (!replaced && !(beta > 1e-14 * std::max(hnorm, 1e-300))) {
                blas_fill_random(ctx, ws.V[n], D, seed ^ (0x5A5A5A5AULL + ++rng_stream), 0);
                replaced = true;
                continue;
            }
            if (beta > 1e-8 * wn) {
                blas_scale_inv(ctx, ws.V[n], ws.V[n], ws.beta(0), D);
                return replaced ? 0.0 : beta;
            }
        }
        EDM_REQUIRE(false, EDM_ENOCONV, "block Lanczos could not orthonormalise a basis vector");
        return 0.0;
    };

    // start block: vec0 (or random) and bs - 1 random vectors, orthonormalised
    if (vec0)
        EDM_HIP_CHECK(hipMemcpyAsync(ws.V[0], vec0, D * sizeof(double2), hipMemcpyHostToDevice, st));
    else
        blas_fill_random(ctx, ws.V[0], D, seed, 0);
    for (int c = 1; c < bs; ++c) blas_fill_random(ctx, ws.V[c], D, seed ^ (0x9E37ULL * c), 0);
    for (int c = 0; c < bs; ++c) {
        std::vector<double2> hs;
        hnorm = 1.0;   // start vectors: cancellation relative to their own norm
        double b0 = careful_column(c, hs);
        EDM_REQUIRE(b0 > 0 || c > 0, EDM_EINVAL, "start vector is zero or not finite");
    }
    hnorm = 0.0;

    std::vector<cd> T((size_t)M * M, cd(0.0, 0.0));
    auto t = [&](int i, int j) -> cd & { return T[(size_t)i * M + j]; };
    std::vector<cd> Bn((size_t)bs * bs, cd(0.0, 0.0));   // coupling of the block after V[M-bs..M)
    std::vector<double> theta, res(M, 0.0);
    std::vector<cd> Y;
    int k = 0, iters = 0, nconv = 0;
    std::vector<const double2 *> xs(bs);
    std::vector<double2 *> ys(bs);

    for (iters = 1; iters <= maxiter; ++iters) {
        for (int j = k; j < M; j += bs) {
            const int p0 = j + bs;
            for (int c = 0; c < bs; ++c) {
                xs[c] = ws.V[j + c];
                ys[c] = ws.V[p0 + c];
            }
            mv(bs, xs.data(), ys.data());
            matvecs += bs;
            // fast path: every column's CGS2 + norm + scaling queued, one read-back
            for (int c = 0; c < bs; ++c) {
                queue_column(ws, c, p0 + c);
                blas_scale_inv(ctx, ws.V[p0 + c], ws.V[p0 + c], ws.beta(c), D);
            }
            std::vector<std::vector<double2>> hs;
            std::vector<double> beta;
            read_columns(ws, bs, p0, hs, beta);
            double est = 0.0;
            for (int c = 0; c < bs; ++c) {
                double col = beta[c];
                for (int r = 0; r < bs; ++r) col += std::hypot(hs[c][j + r].x, hs[c][j + r].y);
                est = std::max(est, col);
            }
            hnorm = std::max(hnorm, est);
            bool clean = true;
            for (int c = 0; c < bs && clean; ++c) {
                double hn = 0.0;
                for (const double2 &v : hs[c]) hn += v.x * v.x + v.y * v.y;
                double wn = std::sqrt(hn + beta[c] * beta[c]);
                clean = beta[c] > 1e-14 * std::max(hnorm, 1e-300) && beta[c] > 1e-8 * wn;
            }
            if (!clean) {
                // slow path: recompute H X and orthonormalise column by column
                mv(bs, xs.data(), ys.data());
                matvecs += bs;
                for (int c = 0; c < bs; ++c) beta[c] = careful_column(p0 + c, hs[c]);
            }
            // diagonal block (Hermitian part) and the triangular coupling to the new block
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    cd a(hs[c][j + r].x, hs[c][j + r].y), b(hs[r][j + c].x, -hs[r][j + c].y);
                    t(j + r, j + c) = r == c ? cd(a.real(), 0.0) : 0.5 * (a + b);
                }
            for (int c = 0; c < bs; ++c)
                for (int r = 0; r < bs; ++r) {
                    cd bv = r < c ? cd(hs[c][p0 + r].x, hs[c][p0 + r].y)
                                  : (r == c ? cd(beta[c], 0.0) : cd(0.0, 0.0));
                    if (p0 < M) {
                        t(p0 + r, j + c) = bv;
                        t(j + c, p0 + r) = std::conj(bv);
                    } else {
                        Bn[(size_t)r * bs + c] = bv;
                    }
                }
        }
        // Rayleigh-Ritz
        std::vector<cd> A = T;
        herm_eigh(M, A, theta, Y);
        nconv = 0;
        for (int i = 0; i < M; ++i) {
            double s2 = 0.0;
            for (int r = 0; r < bs; ++r) {
                cd v = 0.0;
                for (int s = 0; s < bs; ++s) v += Bn[(size_t)r * bs + s] * Y[(size_t)i * M + M - bs + s];
                s2 += std::norm(v);
            }
            res[i] = std::sqrt(s2);
        }
        while (nconv < nev && res[nconv] <= tol) ++nconv;
        if (nconv >= nev || iters == maxiter) break;
        // thick restart: keep about the lowest (3M + 2 nconv) / 5 Ritz vectors (at least
        // nev), rounded down so that M - keep is a whole number q of blocks: rounding up
        // instead can leave one block step per cycle, which converges far more slowly
        int keep = std::max((3 * M + 2 * nconv) / 5, nev);
        int q = std::max(1, (M - keep + bs - 1) / bs);
        while (q > 1 && M - q * bs < nev) --q;
        keep = M - q * bs;
        EDM_HIP_CHECK(hipMemcpyAsync(ws.d_Y, Y.data(), (size_t)M * keep * sizeof(double2),
                                     hipMemcpyHostToDevice, st));
        blas_rotate_c(ctx, (const double2 *const *)ws.d_vptrs, M, ws.d_Y, M, keep,
                      (double2 *const *)ws.d_tmpptrs, D);
        EDM_HIP_CHECK(hipStreamSynchronize(st));
        for (int c = 0; c < keep; ++c) std::swap(ws.V[c], ws.tmp[c]);
        for (int r = 0; r < bs; ++r) std::swap(ws.V[keep + r], ws.V[M + r]);
        ws.upload_ptrs();
        std::fill(T.begin(), T.end(), cd(0.0, 0.0));
        for (int c = 0; c < keep; ++c) {
            t(c, c) = theta[c];
            for (int r = 0; r < bs; ++r) {
                cd s = 0.0;
                for (int u = 0; u < bs; ++u) s += Bn[(size_t)r * bs + u] * Y[(size_t)c * M + M - bs + u];
                t(keep + r, c) = s;
                t(c, keep + r) = std::conj(s);
            }
        }
        k = keep;
    }
    if (iters > maxiter) iters = maxiter;

    for (int i = 0; i < nev; ++i) vals[i] = theta[i];
    double maxres = 0.0;
    for (int i = 0; i < nev; ++i) maxres = std::max(maxres, res[i]);
    if (vecs) {
        EDM_HIP_CHECK(hipMemcpyAsync(ws.d_Y, Y.data(), (size_t)M * nev * sizeof(double2),
                                     hipMemcpyHostToDevice, st));
        blas_rotate_c(ctx, (const double2 *const *)ws.d_vptrs, M, ws.d_Y, M, nev,
                      (double2 *const *)ws.d_tmpptrs, D);
        for (int i = 0; i < nev; ++i)
            EDM_HIP_CHECK(hipMemcpyAsync(vecs + 2 * (size_t)i * D, ws.tmp[i], D * sizeof(double2),
                                         hipMemcpyDeviceToHost, st));
        EDM_HIP_CHECK(hipStreamSynchronize(st));
    }
    if (info) {
        info->iterations = iters;
        info->matvecs = matvecs;
        info->converged = nconv;
        info->krylovdim = M;
        info->max_residual = maxres;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return nconv >= nev ? EDM_OK : EDM_ENOCONV;
}

int run_eigsolve_block_op(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int bs, int nev,
                          int krylovdim, double tol, int maxiter, uint64_t seed,
                          const double *vec0, double *vals, double *vecs, edm_solve_info *info)
{
    const int64_t D = b->D;
    const size_t xb = (size_t)std::min(bs, 8) * (size_t)std::max<int64_t>(D, 1) * sizeof(double2);
    double2 *xi = (double2 *)ctx->alloc(xb);
    int rc;
    try {
        rc = run_eigsolve_block(
            ctx, D, bs,
            [&](int nv, const double2 *const *x, double2 *const *y) {
                launch_matvec_block(ctx, op, b, nv, x, xi, y);
            },
            nev, krylovdim, tol, maxiter, seed, vec0, vals, vecs, info);
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->free(xi, xb);
        throw;
    }
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->free(xi, xb);
    return rc;
}

// ---------------------------------------------------------------- general (non-Hermitian)
// Complex Schur form of a general n x n matrix and the Krylov-Schur iteration built on it,
// for KrylovKit.eigsolve(A, x0, howmany, :SR; ishermitian=false) as krylov_map_solve and
// krylov_matrix_solve call it (MomentumED/src/method/linear_map.jl:179-191,
// sparse_matrix.jl:248-255). The Schur form: Householder reduction to upper Hessenberg form,
// then implicit single-shift QR sweeps with Wilkinson shifts (an exceptional shift after
// every ten sweeps without deflation) until every subdiagonal entry is negligible (Golub &
// Van Loan, 4th ed., Sec. 7.4-7.5, in complex arithmetic); the diagonal is then ordered by
// ascending real part (then imaginary part) by swapping neighbouring diagonal entries with
// one plane rotation each (Sec. 7.6.2). The iteration: Arnoldi with full re-orthogonalisation
// (CGS2) to krylovdim vectors, then a restart that keeps the leading Schur vectors of the
// projected matrix (Stewart, SIAM J. Matrix Anal. Appl. 23, 2001), which keeps the Arnoldi
// relation A V_k = V_k T_k + v b^H exact without implicit shifts.
namespace {

using cplx = std::complex<double>;

// G = [[c, s], [-conj(s), c]], c real, with G [f; g] = [r; 0]
void plane_rotation(cplx f, cplx g, double &c, cplx &s)
{
    const double af = std::abs(f), ag = std::abs(g);
    if (ag == 0.0) { c = 1.0; s = 0.0; return; }
    if (af == 0.0) { c = 0.0; s = std::conj(g) / ag; return; }
    const double nu = std::hypot(af, ag);
    c = af / nu;
    s = (f / af) * std::conj(g) / nu;
}

struct Mat {
    int n;
    std::vector<cplx> a;   // row-major
    cplx &operator()(int i, int j) { return a[(size_t)i * n + j]; }
};

// rows k, k+1 of M <- G [rows], columns [j0, j1)
void rot_rows(Mat &M, int k, double c, cplx s, int j0, int j1)
{
    for (int j = j0; j < j1; ++j) {
        const cplx x = M(k, j), y = M(k + 1, j);
        M(k, j) = c * x + s * y;
        M(k + 1, j) = -std::conj(s) * x + c * y;
    }
}
// columns k, k+1 of M <- [columns] G^H, rows [i0, i1)
void rot_cols(Mat &M, int k, double c, cplx s, int i0, int i1)
{
    for (int i = i0; i < i1; ++i) {
        const cplx x = M(i, k), y = M(i, k + 1);
        M(i, k) = c * x + std::conj(s) * y;
        M(i, k + 1) = -s * x + c * y;
    }
}

bool schur_less(cplx x, cplx y)
{
    return x.real() < y.real() || (x.real() == y.real() && x.imag() < y.imag());
}

// A = Z T Z^H: on return A holds T (upper triangular, diagonal ascending by real part) and Z
// the unitary Schur basis (column c = Schur vector c). False when the QR sweeps do not
// converge.
bool complex_schur(Mat &A, Mat &Z)
{
    const int n = A.n;
    Z.n = n;
    Z.a.assign((size_t)n * n, cplx(0.0, 0.0));
    for (int i = 0; i < n; ++i) Z(i, i) = 1.0;
    double anorm = 0.0;
    for (const cplx &x : A.a) anorm = std::max(anorm, std::abs(x));
    if (anorm == 0.0) return true;
    // work at unit scale (a power of two, exact), so squared norms neither under- nor overflow
    const int ex = scale_exponent(anorm);
    const double down = std::ldexp(1.0, -ex), up = std::ldexp(1.0, ex);
    for (cplx &x : A.a) x *= down;
    anorm *= down;
    struct Rescale {
        Mat &A; double up;
        ~Rescale() { for (cplx &x : A.a) x *= up; }
    } rescale{A, up};
    // Householder reduction to upper Hessenberg form
    std::vector<cplx> v(n), w(n);
    for (int k = 0; k + 2 < n; ++k) {
        double xn = 0.0, tail = 0.0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(A(i, k));
        for (int i = k + 2; i < n; ++i) tail += std::norm(A(i, k));
        if (tail == 0.0) continue;
        xn = std::sqrt(xn);
        const cplx x0 = A(k + 1, k);
        const double ax0 = std::abs(x0);
        const cplx alpha = -(ax0 > 0.0 ? x0 / ax0 : cplx(1.0, 0.0)) * xn;
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += std::norm(v[i]);
        const double beta = 2.0 / vv;
        // A <- H A on rows k+1.., columns k..: A -= beta v (v^H A)
        for (int j = k; j < n; ++j) {
            cplx s = 0.0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * A(i, j);
            s *= beta;
            for (int i = k + 1; i < n; ++i) A(i, j) -= v[i] * s;
        }
        // A <- A H and Z <- Z H on columns k+1..: M -= beta (M v) v^H
        for (Mat *M : {&A, &Z})
            for (int i = 0; i < n; ++i) {
                cplx s = 0.0;
                for (int j = k + 1; j < n; ++j) s += (*M)(i, j) * v[j];
                s *= beta;
                for (int j = k + 1; j < n; ++j) (*M)(i, j) -= s * std::conj(v[j]);
            }
        A(k + 1, k) = alpha;
        for (int i = k + 2; i < n; ++i) A(i, k) = 0.0;
    }
    // implicit single-shift QR sweeps on the active block [l, hi]
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, its = 0, total = 0;
    while (hi > 0) {
        int l = hi;
        for (; l > 0; --l) {
            double sc = std::abs(A(l - 1, l - 1)) + std::abs(A(l, l));
            if (sc == 0.0) sc = anorm;
            if (std::abs(A(l, l - 1)) <= eps * sc) {
                A(l, l - 1) = 0.0;
                break;
            }
        }
        if (l == hi) {
            --hi;
            its = 0;
            continue;
        }
        if (++total > 60 * n) return false;
        cplx mu;
        if (its > 0 && its % 10 == 0) {
            mu = A(hi, hi) + 0.75 * std::abs(A(hi, hi - 1));   // exceptional shift
        } else {
            const cplx a = A(hi - 1, hi - 1), b = A(hi - 1, hi), c = A(hi, hi - 1), d = A(hi, hi);
            const cplx half = 0.5 * (a - d);
            const cplx disc = std::sqrt(half * half + b * c);
            const cplx m1 = 0.5 * (a + d) + disc, m2 = 0.5 * (a + d) - disc;
            mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        }
        ++its;
        for (int k = l; k < hi; ++k) {
            double c;
            cplx s;
            if (k == l) plane_rotation(A(l, l) - mu, A(l + 1, l), c, s);
            else plane_rotation(A(k, k - 1), A(k + 1, k - 1), c, s);
            rot_rows(A, k, c, s, k == l ? l : k - 1, n);
            rot_cols(A, k, c, s, 0, std::min(k + 3, hi + 1));
            rot_cols(Z, k, c, s, 0, n);
            if (k > l) A(k + 1, k - 1) = 0.0;
        }
    }
    for (int i = 1; i < n; ++i)
        for (int j = 0; j < i; ++j) A(i, j) = 0.0;
    // order the diagonal: move the smallest remaining entry to position p by neighbour swaps
    for (int p = 0; p < n; ++p) {
        int q = p;
        for (int r = p + 1; r < n; ++r)
            if (schur_less(A(r, r), A(q, q))) q = r;
        for (int k = q - 1; k >= p; --k) {
            const cplx t11 = A(k, k), t22 = A(k + 1, k + 1);
            double c;
            cplx s;
            plane_rotation(A(k, k + 1), t22 - t11, c, s);
            rot_rows(A, k, c, s, k, n);
            rot_cols(A, k, c, s, 0, k + 2);
            rot_cols(Z, k, c, s, 0, n);
            A(k + 1, k) = 0.0;
            A(k, k) = t22;
            A(k + 1, k + 1) = t11;
        }
    }
    return true;
}

// eigenvector y of the upper triangular T for its i-th diagonal entry (y[j] = 0 for j > i),
// by back substitution; unit norm
std::vector<cplx> tri_eigvec(Mat &T, int i)
{
    const int n = T.n;
    double tnorm = 0.0;
    for (const cplx &x : T.a) tnorm = std::max(tnorm, std::abs(x));
    const double small = std::max(tnorm, 1e-300) * std::numeric_limits<double>::epsilon();
    std::vector<cplx> y(n, cplx(0.0, 0.0));
    y[i] = 1.0;
    const cplx lam = T(i, i);
    for (int j = i - 1; j >= 0; --j) {
        cplx s = 0.0;
        for (int l = j + 1; l <= i; ++l) s += T(j, l) * y[l];
        cplx den = T(j, j) - lam;
        if (std::abs(den) < small) den = small;
        y[j] = -s / den;
    }
    double nrm = 0.0;
    for (const cplx &x : y) nrm += std::norm(x);
    nrm = std::sqrt(nrm);
    for (cplx &x : y) x /= nrm;
    return y;
}

}  // namespace

int projected_schur(int n, const double *A, double *T, double *Z)
{
    Mat a{n, std::vector<cplx>((size_t)n * n)}, z{n, {}};
    for (size_t q = 0; q < a.a.size(); ++q) a.a[q] = cplx(A[2 * q], A[2 * q + 1]);
    EDM_REQUIRE(complex_schur(a, z), EDM_ENOCONV, "Schur decomposition did not converge");
    for (size_t q = 0; q < a.a.size(); ++q) {
        T[2 * q] = a.a[q].real();
        T[2 * q + 1] = a.a[q].imag();
        Z[2 * q] = z.a[q].real();
        Z[2 * q + 1] = z.a[q].imag();
    }
    return EDM_OK;
}

int run_eigsolve_general_mv(edm_ctx *ctx, int64_t D,
                            const std::function<void(const double2 *, double2 *)> &mv, int nev,
                            int krylovdim, double tol, int maxiter, uint64_t seed,
                            const double *vec0, double *vals, double *vecs, edm_solve_info *info)
{
    auto t0 = std::chrono::steady_clock::now();
    EDM_REQUIRE(D > 0, EDM_EINVAL, "empty Hilbert subspace");
    EDM_REQUIRE(nev >= 1, EDM_EINVAL, "N should be at least 1.");
    EDM_REQUIRE(nev <= D, EDM_EINVAL, "more eigenpairs requested than the subspace dimension");
    EDM_REQUIRE(tol > 0, EDM_EINVAL, "tol must be positive");
    EDM_REQUIRE(maxiter >= 1, EDM_EINVAL, "maxiter must be at least 1");
    int m = std::max(krylovdim, nev + 2);
    if ((int64_t)m > D) m = (int)D;
    EDM_REQUIRE(m <= 255, EDM_EINVAL, "krylovdim must be at most 255");
    Workspace ws(ctx, D, m, m);
    {
        int64_t nb = (D + 1023) / 1024;
        ctx->partials((size_t)(nb + 1) * (m + 1) * sizeof(double2));
    }
    const size_t ybytes = (size_t)m * m * sizeof(double2);
    double2 *d_Yc = (double2 *)ctx->alloc(ybytes);
    struct YGuard {
        edm_ctx *c; double2 *p; size_t b;
        ~YGuard() { (void)hipStreamSynchronize(c->stream); c->free(p, b); }
    } yguard{ctx, d_Yc, ybytes};
    hipStream_t st = ctx->stream;
    int matvecs = 0;
    uint64_t rng_stream = 0;
    if (vec0) EDM_HIP_CHECK(hipMemcpyAsync(ws.V[0], vec0, D * sizeof(double2), hipMemcpyHostToDevice, st));
    else blas_fill_random(ctx, ws.V[0], D, seed, 0);
    const double b0 = norm_host(ws, ws.V[0]);
    EDM_REQUIRE(b0 > 0 && std::isfinite(b0), EDM_EINVAL, "start vector is zero or not finite");
    blas_scale_inv(ctx, ws.V[0], ws.V[0], ws.d_beta, D);

    Mat H{m, std::vector<cplx>((size_t)m * m, cplx(0.0, 0.0))};   // the projected matrix
    Mat T{m, {}}, Z{m, {}};
    std::vector<cplx> bres(m);    // residual row: A V Z = V Z T + v_m bres^T
    std::vector<double> res(m, 0.0);
    std::vector<double2> h;
    std::vector<double2> yc((size_t)m * m);
    int k = 0, iters = 0, nconv = 0;
    double beta_last = 0.0, hnorm = 0.0;
    for (iters = 1; iters <= maxiter; ++iters) {
        for (int j = k; j < m; ++j) {
            mv(ws.V[j], ws.V[j + 1]);
            queue_cgs2_norm_scale(ws, j + 1, ws.V[j + 1], 0);
            ++matvecs;
            double beta = 0.0;
            collect_cgs2_norm_scale(ws, j + 1, 0, h, &beta);
            double col = beta;
            for (int r = 0; r <= j; ++r) {
                H(r, j) = cplx(h[r].x, h[r].y);
                col += std::abs(H(r, j));
            }
            hnorm = std::max(hnorm, col);
            if (!(beta > 1e-14 * std::max(hnorm, 1e-300))) {
                // invariant subspace: a fresh random direction orthogonal to the basis
                beta = 0.0;
                for (int attempt = 0; attempt < 4; ++attempt) {
                    blas_fill_random(ctx, ws.V[j + 1], D, seed ^ (0xC3C3C3C3ULL + ++rng_stream), 0);
                    std::vector<double2> hh;
                    cgs2(ws, j + 1, ws.V[j + 1], hh);
                    if (norm_host(ws, ws.V[j + 1]) > 1e-8) break;
                }
                blas_scale_inv(ctx, ws.V[j + 1], ws.V[j + 1], ws.d_beta, D);
            }
            if (j + 1 < m) H(j + 1, j) = beta;
            beta_last = beta;
        }
        T = H;
        EDM_REQUIRE(complex_schur(T, Z), EDM_ENOCONV,
                    "Schur decomposition of the projected matrix did not converge");
        for (int c = 0; c < m; ++c) bres[c] = beta_last * Z(m - 1, c);
        // residual of Ritz pair i: |bres . y_i| for the unit eigenvector y_i of T
        nconv = 0;
        for (int i = 0; i < m; ++i) {
            const std::vector<cplx> y = tri_eigvec(T, i);
            cplx s = 0.0;
            for (int c = 0; c <= i; ++c) s += bres[c] * y[c];
            res[i] = std::abs(s);
        }
        while (nconv < nev && res[nconv] <= tol) ++nconv;
        if (nconv >= nev || iters == maxiter || (int64_t)m == D) break;
        // Krylov-Schur restart: keep the leading Schur vectors
        int keep = (3 * m + 2 * nconv) / 5;
        keep = std::max(keep, nev);
        keep = std::min(keep, m - 1);
        for (int c = 0; c < keep; ++c)
            for (int r = 0; r < m; ++r) yc[(size_t)c * m + r] = make_double2(Z(r, c).real(), Z(r, c).imag());
        EDM_HIP_CHECK(hipMemcpyAsync(d_Yc, yc.data(), (size_t)m * keep * sizeof(double2),
                                     hipMemcpyHostToDevice, st));
        blas_rotate_c(ctx, (const double2 *const *)ws.d_vptrs, m, d_Yc, m, keep,
                      (double2 *const *)ws.d_tmpptrs, D);
        EDM_HIP_CHECK(hipStreamSynchronize(st));
        for (int c = 0; c < keep; ++c) std::swap(ws.V[c], ws.tmp[c]);
        std::swap(ws.V[keep], ws.V[m]);
        ws.upload_ptrs();
        std::fill(H.a.begin(), H.a.end(), cplx(0.0, 0.0));
        for (int r = 0; r < keep; ++r)
            for (int c = r; c < keep; ++c) H(r, c) = T(r, c);
        for (int c = 0; c < keep; ++c) H(keep, c) = bres[c];
        k = keep;
    }
    if (iters > maxiter) iters = maxiter;
    double maxres = 0.0;
    for (int i = 0; i < nev; ++i) {
        vals[2 * i] = T(i, i).real();
        vals[2 * i + 1] = T(i, i).imag();
        maxres = std::max(maxres, res[i]);
    }
    if (vecs) {
        // Ritz vectors V Z y_i
        for (int i = 0; i < nev; ++i) {
            const std::vector<cplx> y = tri_eigvec(T, i);
            for (int r = 0; r < m; ++r) {
                cplx s = 0.0;
                for (int c = 0; c <= i; ++c) s += Z(r, c) * y[c];
                yc[(size_t)i * m + r] = make_double2(s.real(), s.imag());
            }
        }
        EDM_HIP_CHECK(hipMemcpyAsync(d_Yc, yc.data(), (size_t)m * nev * sizeof(double2),
                                     hipMemcpyHostToDevice, st));
        blas_rotate_c(ctx, (const double2 *const *)ws.d_vptrs, m, d_Yc, m, nev,
                      (double2 *const *)ws.d_tmpptrs, D);
        for (int i = 0; i < nev; ++i)
            EDM_HIP_CHECK(hipMemcpyAsync(vecs + 2 * (size_t)i * D, ws.tmp[i], D * sizeof(double2),
                                         hipMemcpyDeviceToHost, st));
        EDM_HIP_CHECK(hipStreamSynchronize(st));
    }
    if (info) {
        info->iterations = iters;
        info->matvecs = matvecs;
        info->converged = nconv;
        info->krylovdim = m;
        info->max_residual = maxres;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return nconv >= nev ? EDM_OK : EDM_ENOCONV;
}
