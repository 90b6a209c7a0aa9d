// edm_enum.hip — momentum-sector basis enumeration on the device.
//
// Replaces the host recursion of ED_momentum_subspaces
// (MomentumED/src/preparation/momentum_decomposition.jl:101-152,188-262) and the colex
// iterator (EDCore/src/colex_mbslist.jl:2-21,32-91). Design:
//   * the states of all conserved components form one mixed-radix rank space; component 0
//     (lowest bits) is the fastest digit, so ascending global rank = ascending uint64
//     (the concatenation order of manybodystate_basis.jl:81-83 with colex = numeric order);
//   * each thread owns R consecutive global ranks: it unranks the first one (colex
//     unranking by the combinatorial number system) and steps with Gosper's
//     next-same-popcount, carrying into the next component on overflow;
//   * the momentum of a state is a sum of per-byte lookup tables staged in LDS, reduced
//     mod Gk (momentum_residue, init_parameter.jl:148-151) to a sector index in
//     momentum_list order (first coordinate fastest);
//   * pass A counts states per (block, sector slot) with LDS atomics; the host scans the
//     block counts; pass B recounts per (thread, slot) in LDS, scans across the block's
//     threads and writes every state at its final position — an order-preserving
//     compaction, so each sector comes out sorted with no sort.
#include "edm_internal.h"

#include <algorithm>

namespace {

constexpr int kEnumBlock = 256;
constexpr int kMaxComp = 8;     // conserved components handled on the device
constexpr int kMaxChunk = 8;    // 8-bit chunks per component (<= 64 bits)
constexpr int kMaxSlots = 64;   // sector slots per counting pass
constexpr int kWriteSlotsOne = 4;   // sector slots per writing pass (one sector)
constexpr int kWriteSlotsAll = 16;  // sector slots per writing pass (all sectors)

struct EnumParams {
    int ncc, n, dim, nchunk;
    int t[kMaxComp];
    uint64_t C[kMaxComp];      // C(n, t_c)
    uint64_t total;            // prod C
    uint64_t R;                // ranks per thread
    int64_t gk[3], kmin[3], krange[3];
    int64_t slot_base;         // sector index of slot 0
    int nslots;
};

__device__ __forceinline__ int64_t dev_residue(int64_t k, int64_t g)
{
    if (g == 0) return k;
    int64_t r = k % g;
    return r < 0 ? r + g : r;
}

// colex unrank: the combination of t elements of [0,n) with 0-based colex rank r
// (colex_mbslist.jl:2-21 with 0-based ranks and bit positions).
__device__ uint64_t colex_unrank(uint64_t r, int t, int n, const uint64_t *binom)
{
    uint64_t s = 0;
    for (int i = t; i >= 1; --i) {
        int c = i - 1;
        while (c + 1 < n && binom[(c + 1) * 65 + i] <= r) ++c;
        s |= 1ULL << c;
        r -= binom[c * 65 + i];
    }
    return s;
}

// next state with the same popcount inside n bits; returns false past the last one
__device__ __forceinline__ bool next_comb(uint64_t &s, int t, int n)
{
    if (t == 0 || t == n) return false;
    uint64_t u = s & (~s + 1ULL);
    uint64_t v = s + u;
    if (v == 0) return false;
    uint64_t nx = v | (((v ^ s) >> 2) / u);
    if (n < 64 && (nx >> n) != 0) return false;
    s = nx;
    return true;
}

__device__ __forceinline__ uint64_t lowest_comb(int t)
{
    return t >= 64 ? ~0ULL : ((1ULL << t) - 1ULL);
}

// sector slot of a full state, or -1
__device__ __forceinline__ int state_slot(const EnumParams &p, const uint64_t *s, const int *mt)
{
    int64_t sums[3] = {0, 0, 0};
    for (int c = 0; c < p.ncc; ++c) {
        uint64_t v = s[c];
        for (int ch = 0; ch < p.nchunk; ++ch) {
            int byte = (int)((v >> (8 * ch)) & 0xFF);
            const int *e = mt + (ch * 256 + byte) * 3;
            sums[0] += e[0];
            sums[1] += e[1];
            sums[2] += e[2];
        }
    }
    int64_t sec = 0, stride = 1;
    for (int d = 0; d < p.dim; ++d) {
        int64_t k = dev_residue(sums[d], p.gk[d]) - p.kmin[d];
        if (k < 0 || k >= p.krange[d]) return -1;
        sec += k * stride;
        stride *= p.krange[d];
    }
    int64_t slot = sec - p.slot_base;
    return (slot >= 0 && slot < p.nslots) ? (int)slot : -1;
}

__device__ __forceinline__ uint64_t concat_state(const EnumParams &p, const uint64_t *s)
{
    uint64_t v = 0;
    for (int c = p.ncc - 1; c >= 0; --c) v = (p.n >= 64 ? 0 : (v << p.n)) | s[c];
    return v;
}

// unrank global rank g into per-component states
__device__ void unrank_global(const EnumParams &p, uint64_t g, uint64_t *s, const uint64_t *binom)
{
    for (int c = 0; c < p.ncc; ++c) {
        uint64_t r = g % p.C[c];
        g /= p.C[c];
        s[c] = colex_unrank(r, p.t[c], p.n, binom);
    }
}

__device__ __forceinline__ void advance(const EnumParams &p, uint64_t *s)
{
    for (int c = 0; c < p.ncc; ++c) {
        if (next_comb(s[c], p.t[c], p.n)) return;
        s[c] = lowest_comb(p.t[c]);
    }
}

__device__ void load_mt(const EnumParams &p, const int *g_mt, int *mt)
{
    int n = p.nchunk * 256 * 3;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mt[i] = g_mt[i];
}

// pass A: per-(block, slot) counts
__global__ void __launch_bounds__(kEnumBlock)
k_enum_count(EnumParams p, const int *__restrict__ g_mt, const uint64_t *__restrict__ binom,
             uint32_t *__restrict__ block_counts)
{
    __shared__ int mt[kMaxChunk * 256 * 3];
    __shared__ uint32_t cnt[kMaxSlots];
    load_mt(p, g_mt, mt);
    for (int i = threadIdx.x; i < p.nslots; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t g0 = tid * p.R;
    if (g0 < p.total) {
        uint64_t nh = p.total - g0 < p.R ? p.total - g0 : p.R;
        uint64_t s[kMaxComp];
        unrank_global(p, g0, s, binom);
        for (uint64_t it = 0; it < nh; ++it) {
            int slot = state_slot(p, s, mt);
            if (slot >= 0) atomicAdd(&cnt[slot], 1u);
            if (it + 1 < nh) advance(p, s);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < p.nslots; i += blockDim.x)
        block_counts[(uint64_t)blockIdx.x * p.nslots + i] = cnt[i];
}

// pass B: order-preserving write. Slot sl's states go to outs[sl]; block_off[blk*nslots+sl]
// = first position in outs[sl] of the block's states of that slot. MAXS bounds nslots (the
// per-thread counts of every slot live in LDS).
template <int MAXS>
__global__ void __launch_bounds__(kEnumBlock)
k_enum_write(EnumParams p, const int *__restrict__ g_mt, const uint64_t *__restrict__ binom,
             const int64_t *__restrict__ block_off, uint64_t *const *__restrict__ outs)
{
    __shared__ int mt[kMaxChunk * 256 * 3];
    __shared__ uint32_t cnt[MAXS * kEnumBlock];
    load_mt(p, g_mt, mt);
    for (int i = threadIdx.x; i < p.nslots * kEnumBlock; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t g0 = tid * p.R;
    uint64_t nh = 0;
    uint64_t s[kMaxComp];
    if (g0 < p.total) {
        nh = p.total - g0 < p.R ? p.total - g0 : p.R;
        unrank_global(p, g0, s, binom);
        for (uint64_t it = 0; it < nh; ++it) {
            int slot = state_slot(p, s, mt);
            if (slot >= 0) cnt[slot * kEnumBlock + threadIdx.x] += 1;
            if (it + 1 < nh) advance(p, s);
        }
    }
    __syncthreads();
    // exclusive scan over the block's threads, one slot per scanning thread
    for (int sl = threadIdx.x; sl < p.nslots; sl += blockDim.x) {
        uint32_t run = 0;
        uint32_t *row = cnt + sl * kEnumBlock;
        for (int t = 0; t < kEnumBlock; ++t) {
            uint32_t v = row[t];
            row[t] = run;
            run += v;
        }
    }
    __syncthreads();
    if (nh) {
        unrank_global(p, g0, s, binom);
        const int64_t *boff = block_off + (uint64_t)blockIdx.x * p.nslots;
        for (uint64_t it = 0; it < nh; ++it) {
            int slot = state_slot(p, s, mt);
            if (slot >= 0) {
                uint32_t &c = cnt[slot * kEnumBlock + threadIdx.x];
                outs[slot][boff[slot] + c] = concat_state(p, s);
                c += 1;
            }
            if (it + 1 < nh) advance(p, s);
        }
    }
}

uint64_t host_binom(int a, int b)
{
    if (b < 0 || b > a) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= b; ++i) r = r * (uint64_t)(a - b + i) / (uint64_t)i;
    return r;
}

void ensure_binom(edm_ctx *ctx)
{
    if (ctx->d_binom) return;
    std::vector<uint64_t> h(65 * 65, 0);
    for (int a = 0; a <= 64; ++a)
        for (int b = 0; b <= 64; ++b) {
            // C(a,b) saturating at UINT64_MAX (never reached for ranks we unrank)
            long double v = 1.0L;
            if (b > a) { h[a * 65 + b] = 0; continue; }
            uint64_t x = 1;
            bool sat = false;
            for (int i = 1; i <= b; ++i) {
                v = v * (a - b + i) / i;
                if (v > 1.8e19L) { sat = true; break; }
                // exact integer update: x * (a-b+i) / i is an integer at every step
                unsigned __int128 t = (unsigned __int128)x * (uint64_t)(a - b + i);
                x = (uint64_t)(t / (uint64_t)i);
            }
            h[a * 65 + b] = sat ? ~0ULL : x;
        }
    ctx->d_binom = (uint64_t *)ctx->alloc(h.size() * sizeof(uint64_t));
    EDM_HIP_CHECK(hipMemcpyAsync(ctx->d_binom, h.data(), h.size() * sizeof(uint64_t),
                                 hipMemcpyHostToDevice, ctx->stream));
}

struct EnumSetup {
    EnumParams p{};
    int *d_mt = nullptr;
    size_t mt_bytes = 0;
    uint64_t nthreads = 0, nblocks = 0;
};

EnumSetup make_setup(edm_ctx *ctx, const SystemSpec &sp)
{
    EDM_REQUIRE(sp.ncc <= kMaxComp, EDM_EINVAL,
                "device enumeration supports at most 8 conserved components");
    EnumSetup s;
    EnumParams &p = s.p;
    p.ncc = sp.ncc;
    p.n = sp.nbits_comp;
    p.dim = sp.dim;
    p.nchunk = (p.n + 7) / 8;
    long double tot = 1.0L;
    p.total = 1;
    for (int c = 0; c < sp.ncc; ++c) {
        p.t[c] = (int)sp.n_each[c];
        p.C[c] = host_binom(p.n, p.t[c]);
        tot *= (long double)p.C[c];
        p.total *= p.C[c];
    }
    EDM_REQUIRE(tot < 1.1e12L, EDM_EINVAL,
                "too many occupation states to enumerate (more than 1.1e12 combinations)");
    for (int d = 0; d < 3; ++d) {
        p.gk[d] = d < sp.dim ? sp.gk[d] : 1;
        p.kmin[d] = d < sp.dim ? sp.kmin[d] : 0;
        p.krange[d] = d < sp.dim ? sp.krange[d] : 1;
    }
    // ranks per thread: enough threads to fill the chip, at most 4096 ranks each
    uint64_t R = p.total / ((uint64_t)kEnumBlock * 8192);
    R = std::max<uint64_t>(16, std::min<uint64_t>(4096, R));
    p.R = R;
    s.nthreads = (p.total + R - 1) / R;
    s.nblocks = (s.nthreads + kEnumBlock - 1) / kEnumBlock;
    // per-byte momentum tables, identical for every component (orbital i of a component
    // has momentum index mod1(i, Nk) because Nk divides the component width)
    std::vector<int> mt((size_t)p.nchunk * 256 * 3, 0);
    for (int ch = 0; ch < p.nchunk; ++ch)
        for (int byte = 0; byte < 256; ++byte)
            for (int bit = 0; bit < 8; ++bit) {
                int b = ch * 8 + bit;
                if (!((byte >> bit) & 1) || b >= p.n) continue;
                for (int d = 0; d < sp.dim; ++d)
                    mt[((size_t)ch * 256 + byte) * 3 + d] += (int)sp.orb_k(b, d);
            }
    s.mt_bytes = mt.size() * sizeof(int);
    s.d_mt = (int *)ctx->alloc(s.mt_bytes);
    EDM_HIP_CHECK(hipMemcpyAsync(s.d_mt, mt.data(), s.mt_bytes, hipMemcpyHostToDevice,
                                 ctx->stream));
    ensure_binom(ctx);
    return s;
}

// counts per (block, slot) for slots [slot_base, slot_base + nslots)
std::vector<uint32_t> run_count(edm_ctx *ctx, EnumSetup &s, int64_t slot_base, int nslots)
{
    s.p.slot_base = slot_base;
    s.p.nslots = nslots;
    size_t bytes = (size_t)s.nblocks * nslots * sizeof(uint32_t);
    uint32_t *d_bc = (uint32_t *)ctx->alloc(bytes);
    hipLaunchKernelGGL(k_enum_count, dim3((unsigned)s.nblocks), dim3(kEnumBlock), 0, ctx->stream,
                       s.p, s.d_mt, ctx->d_binom, d_bc);
    EDM_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> h((size_t)s.nblocks * nslots);
    EDM_HIP_CHECK(hipMemcpyAsync(h.data(), d_bc, bytes, hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->free(d_bc, bytes);
    return h;
}

}  // namespace

void enum_sector_counts(edm_ctx *ctx, const SystemSpec &sp, std::vector<int64_t> &counts)
{
    EnumSetup s = make_setup(ctx, sp);
    counts.assign((size_t)sp.nsec, 0);
    for (int64_t base = 0; base < sp.nsec; base += kMaxSlots) {
        int ns = (int)std::min<int64_t>(kMaxSlots, sp.nsec - base);
        std::vector<uint32_t> h = run_count(ctx, s, base, ns);
        for (uint64_t b = 0; b < s.nblocks; ++b)
            for (int i = 0; i < ns; ++i) counts[base + i] += h[b * ns + i];
    }
    ctx->free(s.d_mt, s.mt_bytes);
}

uint64_t *enum_sector_states(edm_ctx *ctx, const SystemSpec &sp, int64_t sector, int64_t *D_out)
{
    EDM_REQUIRE(sector >= 0 && sector < sp.nsec, EDM_EINVAL, "sector index out of range");
    EnumSetup s = make_setup(ctx, sp);
    uint64_t *d_out = nullptr;
    int64_t *d_off = nullptr;
    uint64_t **d_outs = nullptr;
    size_t obytes = 0;
    try {
        std::vector<uint32_t> h = run_count(ctx, s, sector, 1);
        std::vector<int64_t> off(s.nblocks);
        int64_t run = 0;
        for (uint64_t b = 0; b < s.nblocks; ++b) {
            off[b] = run;
            run += h[b];
        }
        *D_out = run;
        d_out = (uint64_t *)ctx->alloc((size_t)std::max<int64_t>(run, 1) * sizeof(uint64_t));
        obytes = off.size() * sizeof(int64_t);
        d_off = (int64_t *)ctx->alloc(obytes);
        EDM_HIP_CHECK(hipMemcpyAsync(d_off, off.data(), obytes, hipMemcpyHostToDevice, ctx->stream));
        s.p.slot_base = sector;
        s.p.nslots = 1;
        d_outs = (uint64_t **)ctx->alloc(sizeof(uint64_t *));
        EDM_HIP_CHECK(hipMemcpyAsync(d_outs, &d_out, sizeof(uint64_t *), hipMemcpyHostToDevice,
                                     ctx->stream));
        if (run > 0) {
            hipLaunchKernelGGL(k_enum_write<kWriteSlotsOne>, dim3((unsigned)s.nblocks),
                               dim3(kEnumBlock), 0, ctx->stream, s.p, s.d_mt, ctx->d_binom, d_off,
                               d_outs);
            EDM_HIP_CHECK(hipGetLastError());
        }
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        if (d_out) ctx->free(d_out, (size_t)std::max<int64_t>(*D_out, 1) * sizeof(uint64_t));
        if (d_off) ctx->free(d_off, obytes);
        if (d_outs) ctx->free(d_outs, sizeof(uint64_t *));
        ctx->free(s.d_mt, s.mt_bytes);
        throw;
    }
    ctx->free(d_outs, sizeof(uint64_t *));
    ctx->free(d_off, obytes);
    ctx->free(s.d_mt, s.mt_bytes);
    return d_out;
}

void enum_all_sector_states(edm_ctx *ctx, const SystemSpec &sp, std::vector<uint64_t *> &outs,
                            std::vector<int64_t> &D)
{
    EnumSetup s = make_setup(ctx, sp);
    const int64_t nsec = sp.nsec;
    outs.assign((size_t)nsec, nullptr);
    D.assign((size_t)nsec, 0);
    int64_t *d_off = nullptr;
    uint64_t **d_outs = nullptr;
    const size_t obytes = (size_t)s.nblocks * kWriteSlotsAll * sizeof(int64_t);
    auto release_outs = [&]() {
        for (int64_t k = 0; k < nsec; ++k)
            if (outs[k]) ctx->free(outs[k], (size_t)std::max<int64_t>(D[k], 1) * sizeof(uint64_t));
    };
    try {
        // one counting pass per 64 sectors: per-(block, sector) counts, then each sector's
        // block offsets (exclusive scan over blocks) and size
        std::vector<int64_t> off((size_t)s.nblocks * nsec);
        for (int64_t base = 0; base < nsec; base += kMaxSlots) {
            int ns = (int)std::min<int64_t>(kMaxSlots, nsec - base);
            std::vector<uint32_t> h = run_count(ctx, s, base, ns);
            for (int i = 0; i < ns; ++i) {
                int64_t run = 0;
                for (uint64_t b = 0; b < s.nblocks; ++b) {
                    off[b * nsec + base + i] = run;
                    run += h[b * ns + i];
                }
                D[base + i] = run;
            }
        }
        for (int64_t k = 0; k < nsec; ++k)
            outs[k] = (uint64_t *)ctx->alloc((size_t)std::max<int64_t>(D[k], 1) * sizeof(uint64_t));
        d_off = (int64_t *)ctx->alloc(obytes);
        d_outs = (uint64_t **)ctx->alloc(kWriteSlotsAll * sizeof(uint64_t *));
        // one writing pass per 16 sectors
        std::vector<int64_t> hoff((size_t)s.nblocks * kWriteSlotsAll);
        for (int64_t base = 0; base < nsec; base += kWriteSlotsAll) {
            int ns = (int)std::min<int64_t>(kWriteSlotsAll, nsec - base);
            int64_t tot = 0;
            for (int i = 0; i < ns; ++i) tot += D[base + i];
            if (tot == 0) continue;
            for (uint64_t b = 0; b < s.nblocks; ++b)
                for (int i = 0; i < ns; ++i) hoff[b * ns + i] = off[b * nsec + base + i];
            // the previous pass reads d_off / d_outs: wait before overwriting them
            EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            EDM_HIP_CHECK(hipMemcpyAsync(d_off, hoff.data(), (size_t)s.nblocks * ns * sizeof(int64_t),
                                         hipMemcpyHostToDevice, ctx->stream));
            EDM_HIP_CHECK(hipMemcpyAsync(d_outs, outs.data() + base, ns * sizeof(uint64_t *),
                                         hipMemcpyHostToDevice, ctx->stream));
            s.p.slot_base = base;
            s.p.nslots = ns;
            hipLaunchKernelGGL(k_enum_write<kWriteSlotsAll>, dim3((unsigned)s.nblocks),
                               dim3(kEnumBlock), 0, ctx->stream, s.p, s.d_mt, ctx->d_binom, d_off,
                               d_outs);
            EDM_HIP_CHECK(hipGetLastError());
        }
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        release_outs();
        outs.clear();
        if (d_off) ctx->free(d_off, obytes);
        if (d_outs) ctx->free(d_outs, kWriteSlotsAll * sizeof(uint64_t *));
        ctx->free(s.d_mt, s.mt_bytes);
        throw;
    }
    ctx->free(d_off, obytes);
    ctx->free(d_outs, kWriteSlotsAll * sizeof(uint64_t *));
    ctx->free(s.d_mt, s.mt_bytes);
}
