// edm_rank.cpp — two-table (Lin) ranking of a complete momentum sector.
//
// The reference finds a partner state's index by binary search over the sorted basis
// (EDCore/src/types/hilbert_subspace.jl:67-72, device twin gpu_interface.jl:98-113): about
// log2(D) dependent loads per nonzero. For a complete sector (every state with the sector's
// particle numbers and total momentum) the sorted position has a closed form. Split a state
// t into hi = t >> nlo and lo = t & (2^nlo - 1), and let q(.) be the quantum numbers of a bit
// field (particles per conserved component, momentum sums):
//   index(t) = J_hi[hi] + J_lo[lo]
//   J_lo[lo] = #{ lo' < lo : q(lo') == q(lo) }
//   J_hi[hi] = sum_{hi' < hi} #{ lo : q(hi') + q(lo) == Q_sector }
// Both are exactly the number of basis states below t, so the index equals the
// searchsortedfirst position bit for bit. Two independent loads replace ~log2(D) dependent ones.
#include "edm_internal.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <unordered_map>

namespace {

constexpr int kMaxRankBits = 52;

struct QKey {
    uint64_t counts;      // 8 bits per conserved component
    int64_t m[3];
    bool operator==(const QKey &o) const
    {
        return counts == o.counts && m[0] == o.m[0] && m[1] == o.m[1] && m[2] == o.m[2];
    }
};
struct QHash {
    size_t operator()(const QKey &k) const
    {
        uint64_t h = k.counts * 0x9E3779B97F4A7C15ULL;
        for (int d = 0; d < 3; ++d) h = (h ^ (uint64_t)k.m[d]) * 0xBF58476D1CE4E5B9ULL;
        return (size_t)(h ^ (h >> 31));
    }
};

int64_t residue(int64_t k, int64_t g)
{
    if (g == 0) return k;
    int64_t r = k % g;
    return r < 0 ? r + g : r;
}

QKey qkey(const SystemSpec &sp, uint64_t v, int bit0)
{
    QKey q{0, {0, 0, 0}};
    int64_t m[3] = {0, 0, 0};
    while (v) {
        int b = __builtin_ctzll(v) + bit0;
        v &= v - 1;
        q.counts += 1ULL << (8 * sp.comp_of_bit(b));
        for (int d = 0; d < sp.dim; ++d) m[d] += sp.orb_k(b, d);
    }
    for (int d = 0; d < sp.dim; ++d) q.m[d] = residue(m[d], sp.gk[d]);
    return q;
}

// J_lo and the per-quantum-number counts of the low field: the same for every sector of a
// system (they do not involve the sector's momentum)
struct LowTables {
    int nlo = 0, nhi = 0;
    std::vector<uint32_t> jlo;
    std::unordered_map<QKey, uint32_t, QHash> cnt;
};

// Two tables of 2^(nbits/2) entries each (at 52 bits 2 x 256 MB, built on the host by a pass
// over every value of each half) pay for themselves only beside a basis of comparable size:
// they are built when they take at most 64 MB or at most four times the basis' 8 D bytes;
// a small basis of a wide system keeps the bucketed binary search.
static bool rank_tables_worth(int nbits, int64_t D)
{
    if (nbits < 2 || nbits > kMaxRankBits || D <= 0 || D >= (int64_t)0xFFFFFFFFLL) return false;
    const int nlo = nbits / 2, nhi = nbits - nlo;
    const uint64_t bytes = ((1ULL << nlo) + (1ULL << nhi)) * sizeof(uint32_t);
    return bytes <= (64ULL << 20) || bytes <= 4ULL * 8ULL * (uint64_t)D;
}

bool rank_applicable(const edm_basis *b)
{
    if (!b->has_spec || b->sector < 0) return false;
    const SystemSpec &sp = b->spec;
    if (sp.ncc > 8) return false;
    return rank_tables_worth(sp.nbits, b->D);
}

LowTables low_tables(const SystemSpec &sp)
{
    LowTables t;
    t.nlo = sp.nbits / 2;
    t.nhi = sp.nbits - t.nlo;
    const uint64_t nlo_vals = 1ULL << t.nlo;
    t.jlo.resize(nlo_vals);
    t.cnt.reserve(4096);
    for (uint64_t v = 0; v < nlo_vals; ++v) {
        QKey q = qkey(sp, v, 0);
        uint32_t &c = t.cnt[q];
        t.jlo[v] = c;
        ++c;
    }
    return t;
}

// J_hi of one sector; false when the count differs from D (not the complete sector)
bool high_table(const SystemSpec &sp, const LowTables &lo, int64_t sector, int64_t D,
                std::vector<uint32_t> &jhi)
{
    const uint64_t nhi_vals = 1ULL << lo.nhi;
    jhi.assign(nhi_vals, 0);
    // target quantum numbers of the sector
    QKey Q{0, {0, 0, 0}};
    for (int c = 0; c < sp.ncc; ++c) Q.counts += (uint64_t)sp.n_each[c] << (8 * c);
    int64_t K[3] = {0, 0, 0};
    sp.sector_momentum(sector, K);
    uint64_t run = 0;
    for (uint64_t u = 0; u < nhi_vals; ++u) {
        jhi[u] = (uint32_t)run;
        QKey qh = qkey(sp, u, lo.nlo);
        // residual quantum numbers needed from the low field
        QKey need{0, {0, 0, 0}};
        bool ok = true;
        for (int c = 0; c < sp.ncc; ++c) {
            int64_t have = (int64_t)((qh.counts >> (8 * c)) & 0xFF);
            int64_t want = sp.n_each[c] - have;
            if (want < 0) { ok = false; break; }
            need.counts += (uint64_t)want << (8 * c);
        }
        if (!ok) continue;
        for (int d = 0; d < sp.dim; ++d) need.m[d] = residue(K[d] - qh.m[d], sp.gk[d]);
        auto it = lo.cnt.find(need);
        if (it != lo.cnt.end()) run += it->second;
    }
    return (int64_t)run == D;
}

void upload_tables(edm_basis *b, int nlo, int nhi, const std::vector<uint32_t> &jlo,
                   const std::vector<uint32_t> &jhi)
{
    edm_ctx *ctx = b->ctx;
    const uint64_t nlo_vals = 1ULL << nlo, nhi_vals = 1ULL << nhi;
    b->nlo = nlo;
    b->nhi = nhi;
    b->d_jlo = (uint32_t *)ctx->alloc(nlo_vals * sizeof(uint32_t));
    b->d_jhi = (uint32_t *)ctx->alloc(nhi_vals * sizeof(uint32_t));
    EDM_HIP_CHECK(hipMemcpyAsync(b->d_jlo, jlo.data(), nlo_vals * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(b->d_jhi, jhi.data(), nhi_vals * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    uint32_t mx = 0;
    for (uint32_t v : jlo) mx = std::max(mx, v);
    std::vector<uint16_t> j16;   // must outlive the asynchronous copy below
    if (mx <= 0xFFFF) {
        j16.assign(jlo.begin(), jlo.end());
        b->d_jlo16 = (uint16_t *)ctx->alloc(nlo_vals * sizeof(uint16_t));
        EDM_HIP_CHECK(hipMemcpyAsync(b->d_jlo16, j16.data(), nlo_vals * sizeof(uint16_t),
                                     hipMemcpyHostToDevice, ctx->stream));
    }
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace

bool build_rank_tables(edm_basis *b)
{
    if (!rank_applicable(b)) return false;
    LowTables lo = low_tables(b->spec);
    std::vector<uint32_t> jhi;
    if (!high_table(b->spec, lo, b->sector, b->D, jhi)) return false;
    upload_tables(b, lo.nlo, lo.nhi, lo.jlo, jhi);
    return true;
}

std::vector<bool> build_rank_tables_many(const std::vector<edm_basis *> &bs)
{
    std::vector<bool> ok(bs.size(), false);
    if (bs.empty()) return ok;
    // every basis here belongs to one system: J_lo once, J_hi per sector on host threads, for
    // the sectors whose tables are worth their size (rank_applicable)
    std::vector<char> want(bs.size(), 0);
    bool any = false;
    for (size_t k = 0; k < bs.size(); ++k) {
        if (bs[k]->spec.nbits != bs[0]->spec.nbits) return ok;
        want[k] = rank_applicable(bs[k]) ? 1 : 0;
        any = any || want[k];
    }
    if (!any) return ok;
    const SystemSpec &sp = bs[0]->spec;
    LowTables lo = low_tables(sp);
    std::vector<std::vector<uint32_t>> jhi(bs.size());
    std::vector<char> good(bs.size(), 0);
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nth = std::min(bs.size(), hw);
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < bs.size();)
            if (want[k]) good[k] = high_table(sp, lo, bs[k]->sector, bs[k]->D, jhi[k]);
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nth; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    for (size_t k = 0; k < bs.size(); ++k)
        if (good[k]) {
            upload_tables(bs[k], lo.nlo, lo.nhi, lo.jlo, jhi[k]);
            ok[k] = true;
        }
    return ok;
}

bool derive_rank_tables(edm_basis *b, const uint64_t *states)
{
    const int nbits = b->nbits;
    const int64_t D = b->D;
    if (!rank_tables_worth(nbits, D)) return false;
    const int nlo = nbits / 2, nhi = nbits - nlo;
    const uint64_t lomask = (1ULL << nlo) - 1ULL;
    constexpr uint32_t kAbsent = 0xFFFFFFFFu;
    std::vector<uint32_t> jhi(1ULL << nhi, kAbsent), jlo(1ULL << nlo, kAbsent);
    for (int64_t i = 0; i < D; ++i) {
        const uint64_t h = states[i] >> nlo, l = states[i] & lomask;
        if (jhi[h] == kAbsent) jhi[h] = (uint32_t)i;   // sorted: the block's first state
        const uint32_t pos = (uint32_t)(i - jhi[h]);
        if (jlo[l] == kAbsent) jlo[l] = pos;
        else if (jlo[l] != pos) return false;         // blocks differ: no two-table rank
    }
    // entries no state uses are never read for a target inside the basis
    for (uint32_t &v : jhi) if (v == kAbsent) v = 0;
    for (uint32_t &v : jlo) if (v == kAbsent) v = 0;
    upload_tables(b, nlo, nhi, jlo, jhi);
    return true;
}
