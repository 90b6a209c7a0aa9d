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
#include <unordered_map>

namespace {

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

}  // namespace

bool build_rank_tables(edm_basis *b)
{
    if (!b->has_spec || b->sector < 0) return false;
    const SystemSpec &sp = b->spec;
    if (sp.ncc > 8) return false;
    if (b->D >= (int64_t)0xFFFFFFFFLL) return false;
    int nbits = sp.nbits;
    if (nbits > 44 || nbits < 2) return false;
    int nlo = nbits / 2, nhi = nbits - nlo;
    uint64_t nlo_vals = 1ULL << nlo, nhi_vals = 1ULL << nhi;

    std::vector<uint32_t> jlo(nlo_vals), jhi(nhi_vals);
    std::unordered_map<QKey, uint32_t, QHash> cnt;
    cnt.reserve(4096);
    for (uint64_t v = 0; v < nlo_vals; ++v) {
        QKey q = qkey(sp, v, 0);
        uint32_t &c = cnt[q];
        jlo[v] = c;
        ++c;
    }
    // target quantum numbers of the sector
    QKey Q{0, {0, 0, 0}};
    for (int c = 0; c < sp.ncc; ++c) Q.counts += (uint64_t)sp.n_each[c] << (8 * c);
    int64_t K[3] = {0, 0, 0};
    sp.sector_momentum(b->sector, K);
    uint64_t run = 0;
    for (uint64_t u = 0; u < nhi_vals; ++u) {
        jhi[u] = (uint32_t)run;
        QKey qh = qkey(sp, u, nlo);
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
        auto it = cnt.find(need);
        if (it != cnt.end()) run += it->second;
    }
    if ((int64_t)run != b->D) return false;   // not the complete sector

    edm_ctx *ctx = b->ctx;
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
    return true;
}
