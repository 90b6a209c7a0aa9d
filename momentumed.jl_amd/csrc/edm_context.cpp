// edm_context.cpp — context, device memory accounting and system description.
#include "edm_internal.h"

#include <atomic>
#include <cstring>

uint64_t edm_next_id()
{
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

static thread_local std::string g_error;

void edm_set_global_error(const std::string &msg) { g_error = msg; }
const char *edm_global_error() { return g_error.c_str(); }

static int64_t residue(int64_t k, int64_t g)   // init_parameter.jl:148-151
{
    if (g == 0) return k;
    int64_t r = k % g;
    return r < 0 ? r + g : r;
}

SystemSpec SystemSpec::from(const edm_system *s)
{
    EDM_REQUIRE(s != nullptr, EDM_EINVAL, "system description is NULL");
    EDM_REQUIRE(s->dim >= 1 && s->dim <= 3, EDM_EINVAL, "dim must be 1..3");
    EDM_REQUIRE(s->nk >= 1, EDM_EINVAL, "Nk must be positive");
    EDM_REQUIRE(s->nc_mix >= 1 && s->nc_conserve >= 1, EDM_EINVAL,
                "Number of components must be positive");
    EDM_REQUIRE((int64_t)s->nk * s->nc_mix * s->nc_conserve <= 64, EDM_EINVAL,
                "The total orbital number exceeds 64 bits.");   // init_parameter.jl:97
    EDM_REQUIRE(s->k_list && s->gk && s->n_each, EDM_EINVAL, "null array in system description");
    SystemSpec sp;
    sp.dim = s->dim;
    sp.nk = s->nk;
    sp.ncm = s->nc_mix;
    sp.ncc = s->nc_conserve;
    sp.k_list.assign(s->k_list, s->k_list + (size_t)s->dim * s->nk);
    sp.gk.assign(s->gk, s->gk + s->dim);
    sp.n_each.assign(s->n_each, s->n_each + s->nc_conserve);
    sp.nbits_comp = s->nk * s->nc_mix;
    sp.nbits = sp.nbits_comp * s->nc_conserve;
    sp.ntot = 0;
    for (int c = 0; c < sp.ncc; ++c) {
        EDM_REQUIRE(sp.n_each[c] >= 0 && sp.n_each[c] <= sp.nbits_comp, EDM_EINVAL,
                    "Invalid number of electrons in one component");
        sp.ntot += sp.n_each[c];
    }
    for (int d = 0; d < sp.dim; ++d)
        EDM_REQUIRE(sp.gk[d] >= 0, EDM_EINVAL, "Gk must all be non-negative integers.");
    sp.kmin.assign(sp.dim, 0);
    sp.krange.assign(sp.dim, 0);
    sp.nsec = 1;
    for (int d = 0; d < sp.dim; ++d) {   // momentum_decomposition.jl:209-219
        if (sp.gk[d] != 0) {
            sp.kmin[d] = 0;
            sp.krange[d] = sp.gk[d];
        } else {
            int64_t mn = sp.k_list[d], mx = sp.k_list[d];
            for (int k = 1; k < sp.nk; ++k) {
                mn = std::min(mn, sp.k_list[d + sp.dim * k]);
                mx = std::max(mx, sp.k_list[d + sp.dim * k]);
            }
            sp.kmin[d] = mn * sp.ntot;
            sp.krange[d] = mx * sp.ntot - mn * sp.ntot + 1;
        }
        sp.nsec *= sp.krange[d];
    }
    return sp;
}

int64_t SystemSpec::sector_of_sums(const int64_t *sums) const
{
    int64_t sec = 0, stride = 1;
    for (int d = 0; d < dim; ++d) {
        int64_t k = residue(sums[d], gk[d]) - kmin[d];
        if (k < 0 || k >= krange[d]) return -1;
        sec += k * stride;
        stride *= krange[d];
    }
    return sec;
}

void SystemSpec::sector_momentum(int64_t sector, int64_t *k) const
{
    for (int d = 0; d < dim; ++d) {
        k[d] = kmin[d] + sector % krange[d];
        sector /= krange[d];
    }
}

void *edm_ctx::alloc(size_t bytes)
{
    if (bytes == 0) bytes = 16;
    bytes = (bytes + 255) & ~(size_t)255;
    if (budget && used + bytes > budget)
        throw EdmError{EDM_ENOMEM, "HBM budget exceeded: need " + std::to_string(bytes) +
                                       " bytes, " + std::to_string(budget - used) + " left"};
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        throw EdmError{EDM_ENOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                       "): " + hipGetErrorString(e)};
    used += bytes;
    return p;
}

void edm_ctx::free(void *p, size_t bytes)
{
    if (!p) return;
    if (bytes == 0) bytes = 16;
    bytes = (bytes + 255) & ~(size_t)255;
    (void)hipFree(p);
    used = used >= bytes ? used - bytes : 0;
}

double *edm_ctx::partials(size_t bytes)
{
    if (bytes > partials_bytes) {
        // kernels queued on the stream may still read the old buffer
        EDM_HIP_CHECK(hipStreamSynchronize(stream));
        if (d_partials) free(d_partials, partials_bytes);
        d_partials = (double *)alloc(bytes);
        partials_bytes = bytes;
    }
    return d_partials;
}
