// edm_matvec.hip — matrix-free H x in gather form.
//
// Semantics: the CPU map of the reference, MomentumED/src/method/linear_map.jl:100-118 —
// for every basis row i, y_i = sum over full-list terms s of <i|s|j> x_j, where the right
// application state_i * s (EDCore/src/multiplication.jl:55-69) needs out(s) ⊆ state_i, gives
// mid = state_i & ~out, needs mid ∩ in(s) = ∅, target mid | in with sign
// (-1)^popcount(mid & parity) (multiplication.jl:4-6), and a target outside the basis is
// dropped (index_fit, hilbert_subspace.jl:70-72). (The reference's GPU kernel,
// gpu_interface.jl:128-163, left-applies and so computes conj(H) x; this file computes H x.)
//
// Instead of testing every term against every row (the reference loops over the whole
// term list per row), the term list is bucketed by creation mask. A row enumerates only the
// buckets it can feed: one key per occupied orbital (one-body) and one per occupied pair
// (two-body), N + N(N-1)/2 keys. Each bucket lists its terms' annihilation masks, parity
// masks and amplitudes. Partner indices come from the two-table rank of a complete sector
// (edm_rank.cpp) — two independent loads — or from binary search when the basis is not a
// complete sector or the operator does not conserve its quantum numbers. One thread owns one
// row and writes y_i once: no atomics.
#include "edm_internal.h"

#include <climits>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <utility>

namespace {

constexpr int kRowBlock = 256;

struct PairHash {
    size_t operator()(const std::pair<uint64_t, uint64_t> &p) const
    {
        uint64_t h = (p.first ^ (p.second * 0x9E3779B97F4A7C15ULL)) * 0xBF58476D1CE4E5B9ULL;
        return (size_t)(h ^ (h >> 31));
    }
};

struct OpDev {
    const uint32_t *keyptr;
    const uint64_t *tin;
    const uint64_t *tpar;
    const double2 *tamp;
    const int32_t *tterm;
    int has_onebody;
};

struct RankLin {
    const uint32_t *jlo, *jhi;
    int nlo;
    uint64_t lomask;
    const uint16_t *jlo16;   // 16-bit copy of jlo (null when an entry exceeds 65535)
    __device__ __forceinline__ int64_t operator()(uint64_t t) const
    {
        return (int64_t)jhi[t >> nlo] + (int64_t)jlo[t & lomask];
    }
};

// searchsortedfirst over the sorted states, returning -1 for a state outside the basis. With
// a bucket index (start != null) the search is confined to the states whose bits above
// `shift` equal the target's: start[h] is the first index with high part >= h, so the bucket
// of h is [start[h], start[h+1]) and holds t if the basis does.
struct RankSearch {
    const uint64_t *states;
    int64_t D;
    const uint32_t *start;
    int shift;
    uint64_t nbuckets;
    __device__ __forceinline__ int64_t operator()(uint64_t t) const
    {
        int64_t lo = 0, hi = D;
        if (start) {
            const uint64_t h = t >> shift;
            if (h >= nbuckets) return -1;   // above every state of the basis
            lo = start[h];
            hi = start[h + 1];
        }
        const int64_t end = hi;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (states[mid] < t) lo = mid + 1; else hi = mid;
        }
        return (lo < end && states[lo] == t) ? lo : -1;
    }
};

// start[h] = lower_bound(states, h << shift) for h < nbuckets, start[nbuckets] = D
__global__ void __launch_bounds__(256)
k_bucket_starts(int64_t D, const uint64_t *__restrict__ states, int shift, uint64_t nbuckets,
                uint32_t *__restrict__ start)
{
    for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h <= nbuckets;
         h += (uint64_t)gridDim.x * blockDim.x) {
        if (h == nbuckets) {
            start[h] = (uint32_t)D;
            continue;
        }
        const uint64_t t = h << shift;
        int64_t lo = 0, hi = D;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (states[mid] < t) lo = mid + 1; else hi = mid;
        }
        start[h] = (uint32_t)lo;
    }
}

// Visit every (term entry, partner) pair of row state s. f(e, j, neg) for off-diagonal
// contributions, fd(e) for diagonal ones (target = s itself).
template <class Rank, class F, class FD>
__device__ __forceinline__ void visit_row(uint64_t s, const OpDev &op, const Rank &rank, F &&f,
                                          FD &&fd)
{
    auto bucket = [&](int key, uint64_t outm) {
        uint32_t e0 = op.keyptr[key], e1 = op.keyptr[key + 1];
        if (e0 == e1) return;
        uint64_t mid = s & ~outm;
        for (uint32_t e = e0; e < e1; ++e) {
            uint64_t in = op.tin[e];
            if (in == outm) { fd(e); continue; }
            if (mid & in) continue;
            uint64_t t = mid | in;
            int neg = __popcll(mid & op.tpar[e]) & 1;
            int64_t j = rank(t);
            if (j < 0) continue;
            f(e, j, neg);
        }
    };
    uint64_t rem = s;
    while (rem) {
        int b = __builtin_ctzll(rem);
        rem &= rem - 1;
        uint64_t ob = 1ULL << b;
        if (op.has_onebody) bucket(b, ob);
        uint64_t low = s & (ob - 1);
        int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            int a = __builtin_ctzll(low);
            low &= low - 1;
            bucket(kb + a, ob | (1ULL << a));
        }
    }
}

// Flags (*miss = 1) a basis state with an off-diagonal target outside the basis: the same
// candidates as visit_row, each looked up by binary search.
__global__ void __launch_bounds__(kRowBlock)
k_closure_miss(int64_t D, const uint64_t *__restrict__ states, OpDev op, RankSearch rank,
               unsigned int *miss)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint64_t s = states[i];
    bool bad = false;
    auto bucket = [&](int key, uint64_t outm) {
        const uint32_t e0 = op.keyptr[key], e1 = op.keyptr[key + 1];
        const uint64_t mid = s & ~outm;
        for (uint32_t e = e0; e < e1 && !bad; ++e) {
            const uint64_t in = op.tin[e];
            if (in == outm || (mid & in)) continue;
            if (rank(mid | in) < 0) bad = true;
        }
    };
    uint64_t rem = s;
    while (rem && !bad) {
        int b = __builtin_ctzll(rem);
        rem &= rem - 1;
        uint64_t ob = 1ULL << b;
        if (op.has_onebody) bucket(b, ob);
        uint64_t low = s & (ob - 1);
        int kb = 64 + ((b * (b - 1)) >> 1);
        while (low && !bad) {
            int a = __builtin_ctzll(low);
            low &= low - 1;
            bucket(kb + a, ob | (1ULL << a));
        }
    }
    if (bad) atomicOr(miss, 1u);
}

template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_matvec_fast(int64_t D, const uint64_t *__restrict__ states, OpDev op, Rank rank,
              const double2 *__restrict__ x, double2 *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    uint64_t s = states[i];
    double2 xi = x[i];
    double yr = 0.0, yi = 0.0;
    visit_row(
        s, op, rank,
        [&](uint32_t e, int64_t j, int neg) {
            double2 a = op.tamp[e];
            double2 xv = x[j];
            double pr = a.x * xv.x - a.y * xv.y;
            double pi = a.x * xv.y + a.y * xv.x;
            if (neg) { yr -= pr; yi -= pi; } else { yr += pr; yi += pi; }
        },
        [&](uint32_t e) {
            double2 a = op.tamp[e];
            yr += a.x * xi.x - a.y * xi.y;
            yi += a.x * xi.y + a.y * xi.x;
        });
    y[i] = make_double2(yr, yi);
}

// Generic terms (any body count): every full-list term is tried on every row, as the
// reference does; used only for operators the bucketed table cannot hold.
template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_matvec_generic(int64_t D, const uint64_t *__restrict__ states, int64_t S,
                 const double2 *__restrict__ amp, const uint64_t *__restrict__ tout,
                 const uint64_t *__restrict__ tin, const uint64_t *__restrict__ tpar, Rank rank,
                 const double2 *__restrict__ x, double2 *__restrict__ y, int adjoint)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    uint64_t s = states[i];
    double yr = 0.0, yi = 0.0;
    for (int64_t e = 0; e < S; ++e) {
        uint64_t out = adjoint ? tin[e] : tout[e];
        uint64_t in = adjoint ? tout[e] : tin[e];
        if ((s & out) != out) continue;
        double2 a = amp[e];
        if (adjoint) a.y = -a.y;
        int64_t j;
        int neg = 0;
        if (in == out) {
            j = i;
        } else {
            uint64_t mid = s & ~out;
            if (mid & in) continue;
            neg = __popcll(mid & tpar[e]) & 1;
            j = rank(mid | in);
            if (j < 0) continue;
        }
        double2 xv = x[j];
        double pr = a.x * xv.x - a.y * xv.y;
        double pi = a.x * xv.y + a.y * xv.x;
        if (neg) { yr -= pr; yi -= pi; } else { yr += pr; yi += pi; }
    }
    y[i] = make_double2(yr, yi);
}

template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_conn_count(int64_t row0, int64_t nrows, const uint64_t *__restrict__ states, OpDev op,
             Rank rank, int64_t *__restrict__ counts)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int64_t c = 0;
    visit_row(states[row0 + r], op, rank, [&](uint32_t, int64_t, int) { ++c; },
              [&](uint32_t) { ++c; });
    counts[r] = c;
}

template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_conn_fill(int64_t row0, int64_t nrows, const uint64_t *__restrict__ states, OpDev op,
            Rank rank, const int64_t *__restrict__ rowptr, int64_t *__restrict__ cols,
            int32_t *__restrict__ terms, int8_t *__restrict__ negs)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int64_t p = rowptr[r];
    int64_t self = row0 + r;
    visit_row(
        states[self], op, rank,
        [&](uint32_t e, int64_t j, int neg) {
            cols[p] = j;
            terms[p] = op.tterm[e];
            negs[p] = (int8_t)neg;
            ++p;
        },
        [&](uint32_t e) {
            cols[p] = self;
            terms[p] = op.tterm[e];
            negs[p] = 0;
            ++p;
        });
}

// Stored-matrix build: every contribution of row r as one sortable key
// (column << 32 | term << 1 | sign), at rowptr[r]..
template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_csr_keys(int64_t D, const uint64_t *__restrict__ states, OpDev op, Rank rank,
           const int64_t *__restrict__ rowptr, uint64_t *__restrict__ keys)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= D) return;
    int64_t p = rowptr[r];
    visit_row(
        states[r], op, rank,
        [&](uint32_t e, int64_t j, int neg) {
            keys[p++] = ((uint64_t)j << 32) | ((uint64_t)(uint32_t)op.tterm[e] << 1) | (uint64_t)neg;
        },
        [&](uint32_t e) {
            keys[p++] = ((uint64_t)r << 32) | ((uint64_t)(uint32_t)op.tterm[e] << 1);
        });
}

template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_conn_totals(int64_t D, const uint64_t *__restrict__ states, OpDev op, Rank rank,
              unsigned long long *__restrict__ tot)
{
    __shared__ unsigned long long sh[2];
    if (threadIdx.x < 2) sh[threadIdx.x] = 0;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long all = 0, off = 0;
    if (i < D)
        visit_row(states[i], op, rank, [&](uint32_t, int64_t, int) { ++all; ++off; },
                  [&](uint32_t) { ++all; });
    atomicAdd(&sh[0], all);
    atomicAdd(&sh[1], off);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&tot[0], sh[0]);
        atomicAdd(&tot[1], sh[1]);
    }
}

template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_verify_rank(int64_t D, const uint64_t *__restrict__ states, Rank rank,
              unsigned int *__restrict__ bad)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    if (rank(states[i]) != i) atomicAdd(bad, 1u);
}

// ---------------------------------------------------------------- LDS-staged matvec
// The bucket offsets, 16-bit annihilation codes and per-key diagonal amplitudes are copied
// into LDS once per workgroup; a row then reads its candidates from LDS (no L2 traffic for
// the term table) and touches global memory only for the accepted entries' amplitudes, the
// rank tables and the x gather.
constexpr int kLdsBlock = 512;
// LDS a workgroup may use on gfx950 (160 KiB per CU)
constexpr size_t kLdsMaxBytes = 160 * 1024;
constexpr int kNumXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin across them
constexpr int kCuThreads = 2048;   // resident threads per CU (32 wavefronts of 64)

struct CompactDev {
    const uint16_t *keyptr;
    const uint16_t *code;
    const double2 *amp;
    const double2 *diag;
    const double2 *amp_pm;   // position-major amplitudes: [p * nkeys + key]
    int nkeys;
    int noff;
    int off_diag;   // byte offsets in the LDS image
    int off_code;
    int lds_bytes;
    int off_ampidx; // amplitude tail of the image (valid when lds_full > 0)
    int off_amps;
    int lds_full;
};

inline CompactDev compact_dev(const CompactTable &ct)
{
    return CompactDev{ct.d_keyptr, ct.d_code, ct.d_amp, ct.d_diag, ct.d_amp_pm, ct.nkeys, (int)ct.noff,
                      (int)ct.off_diag, (int)ct.off_code, (int)ct.lds_bytes,
                      (int)ct.off_ampidx, (int)ct.off_amps, (int)ct.lds_full};
}

// bits p+1..q-1 (pair p<q) — parity mask of a two-orbital set (scattering.jl:7-18)
__device__ __forceinline__ uint64_t pm_pair(int p, int q)
{
    return (1ULL << q) - (2ULL << p);
}
// bits p+1..63 — parity mask of a single orbital
__device__ __forceinline__ uint64_t pm_one(int p)
{
    return p >= 63 ? 0ULL : (~0ULL << (p + 1));
}

// bit x = parity of the set bits of s below x (prefix XOR of s shifted up by one)
__device__ __forceinline__ uint64_t prefix_parity(uint64_t s)
{
    uint64_t p = s << 1;
    p ^= p << 1;
    p ^= p << 2;
    p ^= p << 4;
    p ^= p << 8;
    p ^= p << 16;
    p ^= p << 32;
    return p;
}

// Consecutive workgroups land on different XCDs; remap so each XCD works through one
// contiguous range of rows (its L2 then holds one neighbourhood of x, not eight).
__device__ __forceinline__ int64_t xcd_block(int swz)
{
    int64_t bid = blockIdx.x;
    if (!swz) return bid;
    int64_t nb = gridDim.x;
    int64_t per = nb / kNumXcd, rem = nb % kNumXcd;
    int64_t xcd = bid % kNumXcd, idx = bid / kNumXcd;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Row block of this workgroup when consecutive dispatches go round-robin to the 8 XCDs: with
// chunk G > 1, every 8 G dispatches cover 8 G consecutive blocks, XCD x taking G consecutive
// ones of them (a bijection on each group of 8 G; a tail shorter than that keeps its order).
__device__ __forceinline__ int64_t xcd_chunk_block(int G)
{
    const int64_t bid = blockIdx.x;
    if (G <= 1) return bid;
    const int64_t grp = (int64_t)kNumXcd * G;
    if (bid >= ((int64_t)gridDim.x / grp) * grp) return bid;
    const int64_t xcd = bid % kNumXcd, idx = bid / kNumXcd;
    return (idx / G) * grp + xcd * G + (idx % G);
}

template <class Rank, bool HighFirst>
__global__ void __launch_bounds__(kLdsBlock)
k_matvec_lds(int64_t D, const uint64_t *__restrict__ states, CompactDev t, Rank rank,
             const double2 *__restrict__ x, double2 *__restrict__ y, int swz)
{
    extern __shared__ uint4 lds_u4[];
    {   // stage the contiguous table image (keyptr | diag | code), 16 bytes per thread per step
        const uint4 *g = (const uint4 *)t.keyptr;
        int n16 = t.lds_bytes / 16;
        for (int k = threadIdx.x; k < n16; k += blockDim.x) lds_u4[k] = g[k];
    }
    __syncthreads();
    const unsigned char *base = (const unsigned char *)lds_u4;
    const uint16_t *s_keyptr = (const uint16_t *)base;
    const double2 *s_diag = (const double2 *)(base + t.off_diag);
    const uint16_t *s_code = (const uint16_t *)(base + t.off_code);

    int64_t i = xcd_block(swz) * (int64_t)blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint64_t s = states[i];
    double yr = 0.0, yi = 0.0, dr = 0.0, di = 0.0;

    auto bucket = [&](int key, uint64_t outm, uint64_t pm_out) {
        double2 dg = s_diag[key];
        dr += dg.x;
        di += dg.y;
        int e0 = s_keyptr[key], e1 = s_keyptr[key + 1];
        uint64_t mid = s & ~outm;
        for (int e = e0; e < e1; ++e) {
            unsigned code = s_code[e];
            int c = code & 0xFF, d = code >> 8;
            uint64_t inm = (1ULL << c) | (d != 0xFF ? (1ULL << d) : 0ULL);
            if (mid & inm) continue;
            uint64_t pm_in = d != 0xFF ? pm_pair(c, d) : pm_one(c);
            int neg = __popcll(mid & (pm_out ^ pm_in)) & 1;
            int64_t j = rank(mid | inm);
            if (j < 0) continue;
            double2 a = t.amp[e];
            double2 xv = x[j];
            double pr = a.x * xv.x - a.y * xv.y;
            double pi = a.x * xv.y + a.y * xv.x;
            if (neg) { yr -= pr; yi -= pi; } else { yr += pr; yi += pi; }
        }
    };
    uint64_t rem = s;
    while (rem) {
        int b = HighFirst ? 63 - __builtin_clzll(rem) : __builtin_ctzll(rem);
        rem ^= 1ULL << b;
        uint64_t ob = 1ULL << b;
        if (b < t.nkeys) bucket(b, ob, pm_one(b));
        uint64_t low = s & (ob - 1ULL);
        int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            int a = HighFirst ? 63 - __builtin_clzll(low) : __builtin_ctzll(low);
            low ^= 1ULL << a;
            int key = kb + a;
            if (key < t.nkeys) bucket(key, ob | (1ULL << a), pm_pair(a, b));
        }
    }
    double2 xi = x[i];
    yr += dr * xi.x - di * xi.y;
    yi += dr * xi.y + di * xi.x;
    y[i] = make_double2(yr, yi);
}

// Two-phase partner lookup: issue() starts the table loads, finish() combines them, so a lane
// can have the lookups of several candidates in flight before it needs any of them.
struct LinPending { uint32_t h, l; };
// Per-row view of the rank tables: a target whose high half equals the row's own reuses the
// row's J_hi entry instead of loading it; J_lo is read from the 16-bit copy when one exists.
struct LinRow {
    const uint32_t *jhi;
    const uint32_t *jlo;
    const uint16_t *jlo16;
    int nlo;
    uint64_t lomask;
    uint64_t hs;
    uint32_t jhs;
};
__device__ __forceinline__ LinRow rank_row(const RankLin &r, uint64_t s)
{
    uint64_t hs = s >> r.nlo;
    return LinRow{r.jhi, r.jlo, r.jlo16, r.nlo, r.lomask, hs, r.jhi[hs]};
}
__device__ __forceinline__ LinPending rank_issue(const LinRow &r, uint64_t t)
{
    uint64_t h = t >> r.nlo;
    uint32_t hv = h == r.hs ? r.jhs : r.jhi[h];
    uint32_t lv = r.jlo16 ? (uint32_t)r.jlo16[t & r.lomask] : r.jlo[t & r.lomask];
    return LinPending{hv, lv};
}
__device__ __forceinline__ int64_t rank_finish(const LinRow &, LinPending p)
{
    return (int64_t)p.h + (int64_t)p.l;
}
__device__ __forceinline__ const RankSearch &rank_row(const RankSearch &r, uint64_t) { return r; }
__device__ __forceinline__ int64_t rank_issue(const RankSearch &r, uint64_t t) { return r(t); }
__device__ __forceinline__ int64_t rank_finish(const RankSearch &, int64_t j) { return j; }

// Flip the sign of a double when neg is 1 (exact: toggles the sign bit).
__device__ __forceinline__ double flip_sign(double v, uint32_t neg)
{
    return __hiloint2double(__double2hiint(v) ^ (int)(neg << 31), __double2loint(v));
}

// y += (ar + i ai) * x as four explicit fused multiply-adds (real part: +ar*x.re, then
// -ai*x.im; imaginary part: +ar*x.im, then +ai*x.re). Spelled out so that every kernel that
// uses it rounds identically, whatever the compiler would otherwise contract.
__device__ __forceinline__ void cmac(double &yr, double &yi, double ar, double ai, double2 x)
{
    yr = __builtin_fma(ar, x.x, yr);
    yr = __builtin_fma(-ai, x.y, yr);
    yi = __builtin_fma(ar, x.y, yi);
    yi = __builtin_fma(ai, x.x, yi);
}

// Lookup traits: a complete sector's rank tables find every partner of a sector-conserving
// operator; the binary search may miss (partner outside a partial basis).
template <class Rank> struct RankMayMiss { static constexpr bool value = true; };
template <> struct RankMayMiss<RankLin> { static constexpr bool value = false; };

// Row lookup for the one-candidate path: 32-bit halves (nlo <= 32 and the high half of a
// state below 2^32 for every basis with rank tables), partner index as a 32-bit value. The
// row's own table words are kept: a partner that shares the row's high (low) half takes the
// row's J_hi (J_lo) word instead of a load (about 18 % of the partners each on the 6x6 N=12
// sector, where a term moves both particles within one half).
struct LinRow32 {
    const uint32_t *jhi, *jlo;
    const uint16_t *jlo16;
    int nlo;
    uint32_t lomask, hs, jhs, ls, jls;
};
__device__ __forceinline__ LinRow32 rank_row32(const RankLin &r, uint64_t s)
{
    const uint32_t hs = (uint32_t)(s >> r.nlo);
    const uint32_t ls = (uint32_t)s & (uint32_t)r.lomask;
    const uint32_t jls = r.jlo16 ? (uint32_t)r.jlo16[ls] : r.jlo[ls];
    return LinRow32{r.jhi, r.jlo, r.jlo16, r.nlo, (uint32_t)r.lomask, hs, r.jhi[hs], ls, jls};
}
__device__ __forceinline__ int64_t rank_find(const LinRow32 &r, uint64_t t)
{
    const uint32_t h = (uint32_t)(t >> r.nlo);
    const uint32_t hv = h == r.hs ? r.jhs : r.jhi[h];
    const uint32_t li = (uint32_t)t & r.lomask;
    const uint32_t lv = li == r.ls ? r.jls : (r.jlo16 ? (uint32_t)r.jlo16[li] : r.jlo[li]);
    return (int64_t)(hv + lv);
}
__device__ __forceinline__ const RankSearch &rank_row32(const RankSearch &r, uint64_t) { return r; }
__device__ __forceinline__ int64_t rank_find(const RankSearch &r, uint64_t t) { return r(t); }

// A bucket's view of the tables: the halves of mid (the row state with the bucket's
// annihilated orbitals removed) and their table words, looked up once per bucket. A partner
// whose created orbitals both lie in one half keeps mid's other half, so that half's word
// comes from the view (about 18 % of the partners for each half on the 6x6 N=12 sector).
struct LinMid32 {
    uint32_t hm, jhm, lm, jlm;
};
__device__ __forceinline__ LinMid32 rank_mid32(const LinRow32 &r, uint64_t mid)
{
    const uint32_t hm = (uint32_t)(mid >> r.nlo), lm = (uint32_t)mid & r.lomask;
    const uint32_t jhm = hm == r.hs ? r.jhs : r.jhi[hm];
    const uint32_t jlm = lm == r.ls ? r.jls : (r.jlo16 ? (uint32_t)r.jlo16[lm] : r.jlo[lm]);
    return LinMid32{hm, jhm, lm, jlm};
}
__device__ __forceinline__ int64_t rank_find(const LinRow32 &r, const LinMid32 &m, uint64_t t)
{
    const uint32_t h = (uint32_t)(t >> r.nlo);
    const uint32_t hv = h == m.hm ? m.jhm : (h == r.hs ? r.jhs : r.jhi[h]);
    const uint32_t li = (uint32_t)t & r.lomask;
    const uint32_t lv = li == m.lm ? m.jlm
                      : (li == r.ls ? r.jls : (r.jlo16 ? (uint32_t)r.jlo16[li] : r.jlo[li]));
    return (int64_t)(hv + lv);
}
struct NoMid {};
__device__ __forceinline__ NoMid rank_mid32(const RankSearch &, uint64_t) { return NoMid{}; }
__device__ __forceinline__ int64_t rank_find(const RankSearch &r, NoMid, uint64_t t) { return r(t); }

// Rank tables split at bit kNarrowLo with J_lo (16-bit, 2^kNarrowLo entries) staged in LDS:
// the low-half lookup, which lanes walking different rows make at scattered addresses (about
// half of the L2 requests of the global-table walk on the 6x6 N=12 sector), becomes an LDS
// read; only J_hi stays in global memory, and lanes that walk the same key of rows sharing
// their high bits mostly read the same J_hi word. sjlo is set inside the kernel, after
// staging. A partner that shares its high part with mid takes mid's word (a partner with the
// row's own high part but not mid's needs a conserved component of several orbitals per
// momentum; the reuse check for it cost more than it saved).
struct RankNarrow {
    const uint32_t *jhi;
    const uint16_t *sjlo;
    int nlo;
    uint32_t lomask;
};
template <> struct RankMayMiss<RankNarrow> { static constexpr bool value = false; };
struct NarrowRow {
    const uint32_t *jhi;
    const uint16_t *sjlo;
    int nlo;
    uint32_t lomask, hs, jhs;
};
struct NarrowMid {
    uint32_t hm, jhm;
};
// t >> nlo for nlo < 32 and a result below 2^32 (every basis with a narrow view): one funnel
// shift of the two 32-bit halves instead of a 64-bit shift
__device__ __forceinline__ uint32_t high_part(uint64_t t, int nlo)
{
    return __builtin_amdgcn_alignbit((uint32_t)(t >> 32), (uint32_t)t, (uint32_t)nlo);
}
__device__ __forceinline__ NarrowRow rank_row32(const RankNarrow &r, uint64_t s)
{
    const uint32_t hs = high_part(s, r.nlo);
    return NarrowRow{r.jhi, r.sjlo, r.nlo, r.lomask, hs, r.jhi[hs]};
}
__device__ __forceinline__ NarrowMid rank_mid32(const NarrowRow &r, uint64_t mid)
{
    const uint32_t hm = high_part(mid, r.nlo);
    return NarrowMid{hm, hm == r.hs ? r.jhs : r.jhi[hm]};
}
__device__ __forceinline__ int64_t rank_find(const NarrowRow &r, const NarrowMid &m, uint64_t t)
{
    const uint32_t h = high_part(t, r.nlo);
    const uint32_t hv = h == m.hm ? m.jhm : r.jhi[h];
    return (int64_t)(hv + (uint32_t)r.sjlo[(uint32_t)t & r.lomask]);
}
// The split at kNarrowLo with J_hi in two levels (bases above kNarrowLo + 24 bits; the
// tables of edm_basis::d_n2_top / d_n2_rows): the high part h = t >> kNarrowLo is cut into
// its top kNarrow2Top bits, which select a row, and m low bits, which index the row. A partner
// that shares h with mid takes mid's word, one that shares the top takes mid's row.
constexpr int kNarrow2Top = 24;
constexpr size_t kNarrow2RowsMax = (size_t)128 << 20;   // bytes of rows at most
struct RankNarrow2 {
    const uint32_t *top;
    const uint32_t *rows;
    const uint16_t *sjlo;
    int m;
    uint32_t lomask;   // (1 << kNarrowLo) - 1, as RankNarrow's
};
template <> struct RankMayMiss<RankNarrow2> { static constexpr bool value = false; };
struct Narrow2Row {
    const uint32_t *top, *rows;
    const uint16_t *sjlo;
    int m;
    uint64_t hs;        // the row's high part
    uint32_t tops, rs, jhs;
};
struct Narrow2Mid {
    uint64_t hm;
    uint32_t topm, rm, jhm;
};
__device__ __forceinline__ Narrow2Row rank_row32(const RankNarrow2 &r, uint64_t s)
{
    const uint64_t hs = s >> kNarrowLo;
    const uint32_t tp = (uint32_t)(hs >> r.m);
    const uint32_t rs = r.top[tp];
    const uint32_t mm = (1u << r.m) - 1u;
    return Narrow2Row{r.top, r.rows, r.sjlo, r.m, hs, tp, rs,
                      r.rows[((size_t)rs << r.m) | ((uint32_t)hs & mm)]};
}
__device__ __forceinline__ Narrow2Mid rank_mid32(const Narrow2Row &r, uint64_t mid)
{
    const uint64_t hm = mid >> kNarrowLo;
    if (hm == r.hs) return Narrow2Mid{hm, r.tops, r.rs, r.jhs};
    const uint32_t tp = (uint32_t)(hm >> r.m);
    const uint32_t rm = tp == r.tops ? r.rs : r.top[tp];
    // mid has one or two particles fewer than a basis state, so its top part may be the top of
    // no state (no row): then no partner shares mid's top or high part, and mid's words are
    // never used; the rows must not be read with the missing row's index
    if (rm == 0xFFFFFFFFu) return Narrow2Mid{hm, tp, rm, 0u};
    const uint32_t mm = (1u << r.m) - 1u;
    return Narrow2Mid{hm, tp, rm, r.rows[((size_t)rm << r.m) | ((uint32_t)hm & mm)]};
}
__device__ __forceinline__ int64_t rank_find(const Narrow2Row &r, const Narrow2Mid &m,
                                             uint64_t t)
{
    const uint64_t h = t >> kNarrowLo;
    uint32_t hv = m.jhm;
    if (h != m.hm) {
        const uint32_t tp = (uint32_t)(h >> r.m);
        const uint32_t rr = tp == m.topm ? m.rm : r.top[tp];
        hv = r.rows[((size_t)rr << r.m) | ((uint32_t)h & ((1u << r.m) - 1u))];
    }
    return (int64_t)(hv + (uint32_t)r.sjlo[(uint32_t)t & ((1u << kNarrowLo) - 1u)]);
}
// the two-level tables read from global memory (checks on every basis state)
struct RankNarrow2Global {
    const uint32_t *top, *rows;
    const uint16_t *jlo;
    int m;
    __device__ __forceinline__ int64_t operator()(uint64_t t) const
    {
        const uint64_t h = t >> kNarrowLo;
        const uint32_t r = top[h >> m];
        if (r == 0xFFFFFFFFu) return -1;
        const uint32_t hv = rows[((size_t)r << m) | (size_t)(h & ((1ULL << m) - 1ULL))];
        return (int64_t)hv + (int64_t)jlo[t & ((1ULL << kNarrowLo) - 1ULL)];
    }
};

// the same tables read from global memory (checks of the split, on every basis state)
struct RankNarrowGlobal {
    const uint32_t *jhi;
    const uint16_t *jlo;
    int nlo;
    uint64_t lomask;
    __device__ __forceinline__ int64_t operator()(uint64_t t) const
    {
        return (int64_t)jhi[t >> nlo] + (int64_t)jlo[t & lomask];
    }
};

// One row, one candidate at a time: pair buckets (key >= 64) hold only two-body entries and
// single-orbital buckets only one-body entries, so each bucket kind has its own loop; the
// per-bucket part of the fermion sign is hoisted; the sign is folded into the amplitude.
// AmpLds: the amplitudes come from the LDS tail of the image (distinct values, with a
// conjugation flag per entry) instead of the global per-entry table; the values are the same
// bits, so the result is too.
// s_waitcnt vmcnt(0): the wavefront's loads into LDS (global_load_lds) have landed. The
// compiler does not order every read of the slot after them by itself.
__device__ __forceinline__ void lds_dma_wait()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt 0, expcnt 7, lgkmcnt 15: only vmcnt waits
    asm volatile("" ::: "memory");
}
// s_waitcnt lgkmcnt(0): the wavefront's LDS reads have returned their data.
__device__ __forceinline__ void lds_read_wait()
{
    __builtin_amdgcn_s_waitcnt(0xC07F);   // vmcnt 63, expcnt 7, lgkmcnt 0: only lgkmcnt waits
    asm volatile("" ::: "memory");
}
// Pipe: each accepted partner's x entry is fetched straight into this lane's 16-byte LDS slot
// (global_load_lds, no registers held while it is in flight) and multiplied in when the lane
// takes its next partner (or at the end of the row), so the gather overlaps the walk to the
// next partner. The products are added in the same order, so y is the same bit for bit.
// wave_slot: the wavefront's 64 slots (wave-uniform), lane_slot: this lane's.
template <class Rank, bool HighFirst, bool AmpLds = false, bool Pipe = false,
          class = std::enable_if_t<!(Pipe && AmpLds)>>
__device__ __forceinline__ double2 row_one(int64_t i, uint64_t s, const unsigned char *base,
                                           const CompactDev &t, const Rank &rank,
                                           const double2 *__restrict__ x, int64_t jmin,
                                           int64_t jmax, int with_diag, bool windowed,
                                           double2 *wave_slot = nullptr,
                                           const double2 *lane_slot = nullptr)
{
    const uint16_t *s_keyptr = (const uint16_t *)base;
    const double2 *s_diag = (const double2 *)(base + t.off_diag);
    const uint16_t *s_code = (const uint16_t *)(base + t.off_code);
    const uint16_t *s_ampidx = (const uint16_t *)(base + t.off_ampidx);
    const double2 *s_amps = (const double2 *)(base + t.off_amps);
    double yr = 0.0, yi = 0.0, dr = 0.0, di = 0.0;
    const auto rrow = rank_row32(rank, s);
    const uint64_t ps = prefix_parity(s);
    bool pend = false;            // Pipe: a partner's x is in flight to the lane's slot
    double par = 0.0, pai = 0.0;  // Pipe: its amplitude
    uint32_t pneg = 0;            // Pipe: and sign

    // pmi: the entry's index in the position-major amplitude table
    auto take = [&](int e, int pmi, const auto &mview, uint64_t target, uint32_t neg) {
        int64_t j = rank_find(rrow, mview, target);
        if (RankMayMiss<Rank>::value && j < 0) return;
        if (windowed && (j < jmin || j >= jmax)) return;
        if constexpr (Pipe) {
            // the previous partner first (its x has been in flight since its take); then this
            // partner's amplitude, loaded into the pending registers only now, so that no copy
            // of a load still in flight is needed; its sign is applied when it is consumed
            if (pend) {
                lds_dma_wait();
                cmac(yr, yi, flip_sign(par, pneg), flip_sign(pai, pneg), *lane_slot);
                lds_read_wait();   // the slot has been read before the next load refills it
            }
            asm volatile("" ::: "memory");
            const double2 a = t.amp_pm[pmi];
            par = a.x;
            pai = a.y;
            pneg = neg;
            __builtin_amdgcn_global_load_lds((const void *)(x + j),
                                             (__attribute__((address_space(3))) void *)wave_slot,
                                             16, 0, 0);
            pend = true;
        } else {
            double2 a;
            uint32_t cj = 0;
            if constexpr (AmpLds) {
                const uint32_t m = s_ampidx[e];
                a = s_amps[m & 0x7FFFu];
                cj = m >> 15;
            } else {
                a = t.amp_pm[pmi];
            }
            double2 xv = x[j];
            cmac(yr, yi, flip_sign(a.x, neg), flip_sign(a.y, neg ^ cj), xv);
        }
    };
    auto enter = [&](int key, int &e0, int &e1) {
        if (with_diag && !Pipe) {   // Pipe: summed after the walk, in the same key order
            double2 dg = s_diag[key];
            dr += dg.x;
            di += dg.y;
        }
        e0 = s_keyptr[key];
        e1 = s_keyptr[key + 1];
    };
    uint64_t rem = s;
    while (rem) {
        int b = HighFirst ? 63 - __builtin_clzll(rem) : __builtin_ctzll(rem);
        rem ^= 1ULL << b;
        const uint64_t ob = 1ULL << b;
        if (b < t.nkeys) {   // one-body entries: in = {c}
            int e0, e1;
            enter(b, e0, e1);
            const uint64_t mid = s & ~ob;
            const uint32_t pout = __popcll(mid & pm_one(b)) & 1u;
            const auto mview = rank_mid32(rrow, mid);
            for (int e = e0; e < e1; ++e) {
                const int c = s_code[e] & 0xFF;
                const uint64_t inm = 1ULL << c;
                if (mid & inm) continue;
                const uint32_t neg = (pout + (uint32_t)__popcll(mid & pm_one(c))) & 1u;
                take(e, (e - e0) * t.nkeys + b, mview, mid | inm, neg);
            }
        }
        uint64_t low = s & (ob - 1ULL);
        const int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            int a = HighFirst ? 63 - __builtin_clzll(low) : __builtin_ctzll(low);
            low ^= 1ULL << a;
            const int key = kb + a;
            if (key >= t.nkeys) continue;
            int e0, e1;
            enter(key, e0, e1);
            const uint64_t mid = s & ~(ob | (1ULL << a));
            const uint32_t pout = __popcll(mid & pm_pair(a, b)) & 1u;
            const auto mview = rank_mid32(rrow, mid);
            // bit x: parity of mid's orbitals below x; for free c < d the orbitals of mid
            // strictly between them have parity bit c ^ bit d
            const uint64_t pmid = ps ^ pm_one(a) ^ pm_one(b);
            const int n = e1 - e0;
            int pmi = key;   // position-major index of entry e0 + q
            for (int q = 0; q < n; ++q, pmi += t.nkeys) {   // two-body entries: in = {c, d}, c < d
                const unsigned code = s_code[e0 + q];
                const int c = code & 0xFF, d = code >> 8;
                if (((uint32_t)(mid >> c) | (uint32_t)(mid >> d)) & 1u) continue;
                const uint32_t neg = (pout ^ (uint32_t)(pmid >> c) ^ (uint32_t)(pmid >> d)) & 1u;
                take(e0 + q, pmi, mview, mid | (1ULL << c) | (1ULL << d), neg);
            }
        }
    }
    if constexpr (Pipe) {
        if (pend) {
            lds_dma_wait();
            cmac(yr, yi, flip_sign(par, pneg), flip_sign(pai, pneg), *lane_slot);
        }
        // the diagonal terms of the keys the walk entered, in its order, from the global copy
        // of the table: a load inside the walk would have to wait for the x in flight
        if (with_diag) {
            uint64_t r2 = s;
            while (r2) {
                const int b = HighFirst ? 63 - __builtin_clzll(r2) : __builtin_ctzll(r2);
                r2 ^= 1ULL << b;
                if (b < t.nkeys) {
                    const double2 dg = t.diag[b];
                    dr += dg.x;
                    di += dg.y;
                }
                uint64_t low = s & ((1ULL << b) - 1ULL);
                const int kb = 64 + ((b * (b - 1)) >> 1);
                while (low) {
                    const int a = HighFirst ? 63 - __builtin_clzll(low) : __builtin_ctzll(low);
                    low ^= 1ULL << a;
                    if (kb + a >= t.nkeys) continue;
                    const double2 dg = t.diag[kb + a];
                    dr += dg.x;
                    di += dg.y;
                }
            }
        }
    }
    if (with_diag) cmac(yr, yi, dr, di, x[i]);
    return make_double2(yr, yi);
}

// One row shared by G lanes (G divides 64; the lanes are consecutive in one wavefront). The
// row's keys are walked as by row_one; within a bucket lane g resolves entries e0 + g,
// e0 + g + G, ... (test, sign, partner lookup, amplitude and x loads) and parks its signed
// amplitude and x value in its LDS slot (zeros when the entry does not apply); the row's
// first lane then adds the G slots in entry order with the same cmac as row_one. Adding a
// zero contribution leaves the sum's bits unchanged (it starts at +0 and a sum that starts at
// +0 never becomes -0), so y is bitwise the row_one result for every G. slot: this lane's
// 2 x double2 in LDS, row0: the row's first lane's.
template <class Rank, int G>
__device__ __forceinline__ double2 row_split(int64_t i, uint64_t s, const unsigned char *base,
                                             const CompactDev &t, const Rank &rank,
                                             const double2 *__restrict__ x, int g,
                                             double2 *slot, const double2 *row0)
{
    const uint16_t *s_keyptr = (const uint16_t *)base;
    const double2 *s_diag = (const double2 *)(base + t.off_diag);
    const uint16_t *s_code = (const uint16_t *)(base + t.off_code);
    double yr = 0.0, yi = 0.0, dr = 0.0, di = 0.0;
    const auto rrow = rank_row32(rank, s);

    auto park = [&](int e, bool ok, uint64_t target, uint32_t neg) {
        double2 a = make_double2(0.0, 0.0), xv = make_double2(0.0, 0.0);
        if (ok) {
            const int64_t j = rank_find(rrow, target);
            if (!(RankMayMiss<Rank>::value && j < 0)) {
                const double2 am = t.amp[e];
                a = make_double2(flip_sign(am.x, neg), flip_sign(am.y, neg));
                xv = x[j];
            }
        }
        slot[0] = a;
        slot[1] = xv;
        __builtin_amdgcn_wave_barrier();
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < G; ++q) cmac(yr, yi, row0[2 * q].x, row0[2 * q].y, row0[2 * q + 1]);
        }
        __builtin_amdgcn_wave_barrier();
    };
    auto enter = [&](int key, int &e0, int &e1) {
        const double2 dg = s_diag[key];
        dr += dg.x;
        di += dg.y;
        e0 = s_keyptr[key];
        e1 = s_keyptr[key + 1];
    };
    uint64_t rem = s;
    while (rem) {
        const int b = 63 - __builtin_clzll(rem);
        rem ^= 1ULL << b;
        const uint64_t ob = 1ULL << b;
        if (b < t.nkeys) {   // one-body entries: in = {c}
            int e0, e1;
            enter(b, e0, e1);
            const uint64_t mid = s & ~ob;
            const uint32_t pout = __popcll(mid & pm_one(b)) & 1u;
            for (int c0 = e0; c0 < e1; c0 += G) {
                const int e = c0 + g;
                bool ok = false;
                uint64_t target = 0;
                uint32_t neg = 0;
                if (e < e1) {
                    const int c = s_code[e] & 0xFF;
                    const uint64_t inm = 1ULL << c;
                    ok = !(mid & inm);
                    target = mid | inm;
                    neg = (pout + (uint32_t)__popcll(mid & pm_one(c))) & 1u;
                }
                park(e, ok, target, neg);
            }
        }
        uint64_t low = s & (ob - 1ULL);
        const int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            const int a = 63 - __builtin_clzll(low);
            low ^= 1ULL << a;
            const int key = kb + a;
            if (key >= t.nkeys) continue;
            int e0, e1;
            enter(key, e0, e1);
            const uint64_t mid = s & ~(ob | (1ULL << a));
            const uint32_t pout = __popcll(mid & pm_pair(a, b)) & 1u;
            for (int c0 = e0; c0 < e1; c0 += G) {   // two-body entries: in = {c, d}, c < d
                const int e = c0 + g;
                bool ok = false;
                uint64_t target = 0;
                uint32_t neg = 0;
                if (e < e1) {
                    const unsigned code = s_code[e];
                    const int c = code & 0xFF, d = code >> 8;
                    const uint64_t bd = 1ULL << d;
                    const uint64_t inm = bd | (1ULL << c);
                    ok = !(mid & inm);
                    target = mid | inm;
                    neg = (pout + (uint32_t)__popcll(mid & (bd - (2ULL << c)))) & 1u;
                }
                park(e, ok, target, neg);
            }
        }
    }
    cmac(yr, yi, dr, di, x[i]);
    return make_double2(yr, yi);
}

// One row of the LDS-staged matvec: y_i (with the diagonal when with_diag) over partners in
// [jmin, jmax) (jmin = 0, jmax = INT64_MAX: no window). The term image must already be in
// LDS at `base`.
template <class Rank, bool HighFirst, int U>
__device__ __forceinline__ double2 ldsu_row(int64_t i, uint64_t s, const unsigned char *base,
                                            const CompactDev &t, const Rank &rank,
                                            const double2 *__restrict__ x, int64_t jmin,
                                            int64_t jmax, int with_diag)
{
    if constexpr (U == 1)
        return row_one<Rank, HighFirst>(i, s, base, t, rank, x, jmin, jmax, with_diag,
                                        jmin > 0 || jmax < (int64_t)INT64_MAX);
    const uint16_t *s_keyptr = (const uint16_t *)base;
    const double2 *s_diag = (const double2 *)(base + t.off_diag);
    const uint16_t *s_code = (const uint16_t *)(base + t.off_code);
    double yr = 0.0, yi = 0.0, dr = 0.0, di = 0.0;
    const auto rrow = rank_row(rank, s);
    using Pend = decltype(rank_issue(rrow, 0ULL));

    auto bucket = [&](int key, uint64_t outm, uint64_t pm_out) {
        if (with_diag) {
            double2 dg = s_diag[key];
            dr += dg.x;
            di += dg.y;
        }
        int e0 = s_keyptr[key], e1 = s_keyptr[key + 1];
        uint64_t mid = s & ~outm;
        for (int e = e0; e < e1; e += U) {
            Pend pend[U];
            bool live[U];
            int neg[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                live[u] = false;
                neg[u] = 0;
                if (e + u < e1) {
                    unsigned code = s_code[e + u];
                    int c = code & 0xFF, d = code >> 8;
                    uint64_t inm = (1ULL << c) | (d != 0xFF ? (1ULL << d) : 0ULL);
                    if (!(mid & inm)) {
                        uint64_t pm_in = d != 0xFF ? pm_pair(c, d) : pm_one(c);
                        neg[u] = __popcll(mid & (pm_out ^ pm_in)) & 1;
                        pend[u] = rank_issue(rrow, mid | inm);
                        live[u] = true;
                    }
                }
            }
            int64_t j[U];
            double2 a[U], xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                j[u] = live[u] ? rank_finish(rrow, pend[u]) : -1;
                live[u] = live[u] && j[u] >= jmin && j[u] < jmax;
                if (live[u]) {
                    a[u] = t.amp[e + u];
                    xv[u] = x[j[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (live[u]) {
                    double pr = a[u].x * xv[u].x - a[u].y * xv[u].y;
                    double pi = a[u].x * xv[u].y + a[u].y * xv[u].x;
                    if (neg[u]) { yr -= pr; yi -= pi; } else { yr += pr; yi += pi; }
                }
            }
        }
    };
    uint64_t rem = s;
    while (rem) {
        int b = HighFirst ? 63 - __builtin_clzll(rem) : __builtin_ctzll(rem);
        rem ^= 1ULL << b;
        uint64_t ob = 1ULL << b;
        if (b < t.nkeys) bucket(b, ob, pm_one(b));
        uint64_t low = s & (ob - 1ULL);
        int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            int a = HighFirst ? 63 - __builtin_clzll(low) : __builtin_ctzll(low);
            low ^= 1ULL << a;
            int key = kb + a;
            if (key < t.nkeys) bucket(key, ob | (1ULL << a), pm_pair(a, b));
        }
    }
    if (with_diag) {
        double2 xi = x[i];
        yr += dr * xi.x - di * xi.y;
        yi += dr * xi.y + di * xi.x;
    }
    return make_double2(yr, yi);
}

__device__ __forceinline__ void stage_image(const CompactDev &t, uint4 *lds)
{
    const uint4 *g = (const uint4 *)t.keyptr;
    int n16 = t.lds_bytes / 16;
    for (int k = threadIdx.x; k < n16; k += blockDim.x) lds[k] = g[k];
}

// The image without its diagonal table (keyptr | code), for kernels that read the diagonal
// from global memory: image_nodiag gives the offsets of that layout.
__host__ __device__ __forceinline__ CompactDev image_nodiag(CompactDev t)
{
    t.lds_bytes = t.off_diag + (t.lds_bytes - t.off_code);
    t.off_code = t.off_diag;
    return t;
}
__device__ __forceinline__ void stage_image_nodiag(const CompactDev &t, uint4 *lds)
{
    const uint4 *g = (const uint4 *)t.keyptr;
    const int nk16 = t.off_diag / 16, skip = (t.off_code - t.off_diag) / 16;
    const int n16 = (t.lds_bytes - (t.off_code - t.off_diag)) / 16;
    for (int k = threadIdx.x; k < n16; k += blockDim.x) lds[k] = g[k < nk16 ? k : k + skip];
}

// the image with its amplitude tail
__device__ __forceinline__ void stage_image_full(const CompactDev &t, uint4 *lds)
{
    const uint4 *g = (const uint4 *)t.keyptr;
    int n16 = t.lds_full / 16;
    for (int k = threadIdx.x; k < n16; k += blockDim.x) lds[k] = g[k];
}

// Same traversal as k_matvec_lds, with U candidates of a bucket resolved together (U table
// lookups in flight, then U gathers in flight), a per-row rank view, and an optional partner
// window: only partners j in [jmin, jmax) are gathered, and the result is added to y when
// `accumulate` is set. A matvec split into passes over windows of x keeps each pass's gather
// footprint small.
template <class Rank, bool HighFirst, int U, int BLOCK = kLdsBlock>
__global__ void __launch_bounds__(BLOCK)
k_matvec_ldsu(int64_t D, const uint64_t *__restrict__ states, CompactDev t, Rank rank,
              const double2 *__restrict__ x, double2 *__restrict__ y, int64_t jmin, int64_t jmax,
              int accumulate, int with_diag, int chunk)
{
    extern __shared__ uint4 lds_u4[];
    stage_image(t, lds_u4);
    __syncthreads();
    int64_t i = xcd_chunk_block(chunk) * blockDim.x + threadIdx.x;
    if (i >= D) return;
    double2 v = ldsu_row<Rank, HighFirst, U>(i, states[i], (const unsigned char *)lds_u4, t, rank,
                                             x, jmin, jmax, with_diag);
    if (accumulate) {
        double2 o = y[i];
        v.x += o.x;
        v.y += o.y;
    }
    y[i] = v;
}

// The default matvec with G lanes per row (row_split): kLdsBlock / G rows per workgroup, the
// term image in LDS followed by one 32-byte exchange slot per lane. y is bitwise the one-lane
// kernel's; more lanes per row shorten each thread's walk when there are too few rows to fill
// the device.
template <class Rank, int G>
__global__ void __launch_bounds__(kLdsBlock)
k_matvec_split(int64_t D, const uint64_t *__restrict__ states, CompactDev t, Rank rank,
               const double2 *__restrict__ x, double2 *__restrict__ y)
{
    static_assert(G >= 2 && G <= 64 && (64 % G) == 0, "lanes per row must divide 64");
    extern __shared__ uint4 lds_u4[];
    stage_image(t, lds_u4);
    double2 *slots = (double2 *)((unsigned char *)lds_u4 + ((t.lds_bytes + 15) & ~15));
    __syncthreads();
    const int64_t i = ((int64_t)blockIdx.x * kLdsBlock + threadIdx.x) / G;
    if (i >= D) return;   // all G lanes of a row leave together; no barrier follows
    const int g = (int)(threadIdx.x % G);
    double2 *slot = slots + 2 * threadIdx.x;
    const double2 *row0 = slots + 2 * (threadIdx.x - g);
    const double2 v = row_split<Rank, G>(i, states[i], (const unsigned char *)lds_u4, t, rank, x,
                                         g, slot, row0);
    if (g == 0) y[i] = v;
}

// k_matvec_ldsu<Rank, true, 1, BLOCK> with the amplitudes staged in LDS as well (the whole
// image, lds_full bytes): per accepted candidate one 16-byte global load fewer. y is bitwise
// the same.
template <class Rank, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_matvec_ldsa(int64_t D, const uint64_t *__restrict__ states, CompactDev t, Rank rank,
              const double2 *__restrict__ x, double2 *__restrict__ y)
{
    extern __shared__ uint4 lds_u4[];
    stage_image_full(t, lds_u4);
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= D) return;
    y[i] = row_one<Rank, true, true>(i, states[i], (const unsigned char *)lds_u4, t, rank, x, 0,
                                     INT64_MAX, 1, false);
}

// k_matvec_ldsu's row walk (row_one, same order and arithmetic, so the same y bit for bit)
// with the partner lookup of RankNarrow: the workgroup stages the term image and, 16-byte
// aligned after it, the 2^kNarrowLo-entry J_lo. BLOCK = 1024 keeps the occupancy of the
// 512-row kernel (two workgroups of the ~67 KiB image per CU on the 6x6 N=12 sector).
// the 2^nlo-entry J_lo, 16 bytes per thread per step, 16-byte aligned after the term image;
// returns its LDS address (valid after the barrier that follows)
template <class R>
__device__ __forceinline__ const uint16_t *stage_narrow(const CompactDev &t, const R &r,
                                                        const uint16_t *__restrict__ jlo,
                                                        uint4 *lds)
{
    const int off = (t.lds_bytes + 15) & ~15;
    const uint4 *g = (const uint4 *)jlo;
    uint4 *d = (uint4 *)((unsigned char *)lds + off);
    const int n16 = (int)(((size_t)r.lomask + 1) * sizeof(uint16_t) / 16);
    for (int k = threadIdx.x; k < n16; k += blockDim.x) d[k] = g[k];
    return (const uint16_t *)((const unsigned char *)lds + off);
}

template <int BLOCK, bool PIPE = false, class R = RankNarrow>
__global__ void __launch_bounds__(BLOCK)
k_matvec_narrow(int64_t D, const uint64_t *__restrict__ states, CompactDev t, R rank,
                const uint16_t *__restrict__ jlo, const double2 *__restrict__ x,
                double2 *__restrict__ y, int chunk)
{
    extern __shared__ uint4 lds_u4[];
    __shared__ double2 pipe_slots[PIPE ? BLOCK : 1];
    if constexpr (PIPE) {
        stage_image_nodiag(t, lds_u4);
        t = image_nodiag(t);
    } else {
        stage_image(t, lds_u4);
    }
    const uint16_t *sjlo = stage_narrow(t, rank, jlo, lds_u4);
    __syncthreads();
    rank.sjlo = sjlo;
    const int64_t i = xcd_chunk_block(chunk) * BLOCK + threadIdx.x;
    if (i >= D) return;
    if constexpr (PIPE)
        y[i] = row_one<R, true, false, true>(
            i, states[i], (const unsigned char *)lds_u4, t, rank, x, 0, INT64_MAX, 1, false,
            pipe_slots + (threadIdx.x & ~63u), pipe_slots + threadIdx.x);
    else
        y[i] = row_one<R, true>(i, states[i], (const unsigned char *)lds_u4, t, rank, x, 0,
                                INT64_MAX, 1, false);
}

// J_hi_n[h] = index of the first state whose high part (bits >= nlo) is h (entries of absent
// high parts stay as filled)
__global__ void __launch_bounds__(256)
k_narrow_hi(int64_t D, const uint64_t *__restrict__ states, int nlo, uint32_t *__restrict__ jhi)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint64_t h = states[i] >> nlo;
    if (i == 0 || (states[i - 1] >> nlo) != h) jhi[h] = (uint32_t)i;
}

// J_lo_n[l] = position of state i in its block, for l its low part. In a basis with the
// two-table structure every block that holds l puts it at the same position, so concurrent
// writers store the same value; otherwise the check on every state that follows fails.
// Positions above 65535 are counted in *bad.
__global__ void __launch_bounds__(256)
k_narrow_lo(int64_t D, const uint64_t *__restrict__ states, int nlo,
            const uint32_t *__restrict__ jhi, uint16_t *__restrict__ jlo,
            unsigned int *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint64_t s = states[i];
    const int64_t pos = i - (int64_t)jhi[s >> nlo];
    if (pos < 0 || pos > 0xFFFF) {
        atomicAdd(bad, 1u);
        return;
    }
    jlo[s & ((1ULL << nlo) - 1ULL)] = (uint16_t)pos;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *__restrict__ p, int64_t n, uint32_t v)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = v;
}

// ---------------------------------------------------------------- block matvec
// Y = H X for NV vectors at once. The candidate generation, sign and partner lookup of a row
// are done once and serve all NV vectors; the gathers read x_0..x_{NV-1} of partner j from one
// interleaved copy Xi[j*NV + v] (NV*16 contiguous bytes), so one partner costs one cache
// line for all vectors. Per vector the arithmetic (cmac) and the summation order are those
// of row_one, so column v of Y equals the single-vector matvec of x_v bit for bit.
template <int NV> struct VecPtrs { double2 *p[NV]; };
template <int NV> struct CVecPtrs { const double2 *p[NV]; };

template <class Rank, int NV>
__device__ __forceinline__ void row_block(uint64_t s, const unsigned char *base,
                                          const CompactDev &t, const Rank &rank,
                                          const double2 *__restrict__ xi, double2 *acc,
                                          double2 *dsum)
{
    const uint16_t *s_keyptr = (const uint16_t *)base;
    const double2 *s_diag = (const double2 *)(base + t.off_diag);
    const uint16_t *s_code = (const uint16_t *)(base + t.off_code);
    double dr = 0.0, di = 0.0;
    const auto rrow = rank_row32(rank, s);

    // pmi: the entry's index in the position-major amplitude table
    auto take = [&](int pmi, const auto &mview, uint64_t target, uint32_t neg) {
        int64_t j = rank_find(rrow, mview, target);
        if (RankMayMiss<Rank>::value && j < 0) return;
        double2 a = t.amp_pm[pmi];
        const double2 *xj = xi + j * NV;
        double2 xv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) xv[v] = xj[v];
        double ax = flip_sign(a.x, neg), ay = flip_sign(a.y, neg);
#pragma unroll
        for (int v = 0; v < NV; ++v) cmac(acc[v].x, acc[v].y, ax, ay, xv[v]);
    };
    auto enter = [&](int key, int &e0, int &e1) {
        double2 dg = s_diag[key];
        dr += dg.x;
        di += dg.y;
        e0 = s_keyptr[key];
        e1 = s_keyptr[key + 1];
    };
    uint64_t rem = s;
    while (rem) {
        int b = 63 - __builtin_clzll(rem);
        rem ^= 1ULL << b;
        const uint64_t ob = 1ULL << b;
        if (b < t.nkeys) {
            int e0, e1;
            enter(b, e0, e1);
            const uint64_t mid = s & ~ob;
            const uint32_t pout = __popcll(mid & pm_one(b)) & 1u;
            const auto mview = rank_mid32(rrow, mid);
            for (int e = e0; e < e1; ++e) {
                const int c = s_code[e] & 0xFF;
                const uint64_t inm = 1ULL << c;
                if (mid & inm) continue;
                const uint32_t neg = (pout + (uint32_t)__popcll(mid & pm_one(c))) & 1u;
                take((e - e0) * t.nkeys + b, mview, mid | inm, neg);
            }
        }
        uint64_t low = s & (ob - 1ULL);
        const int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            int a = 63 - __builtin_clzll(low);
            low ^= 1ULL << a;
            const int key = kb + a;
            if (key >= t.nkeys) continue;
            int e0, e1;
            enter(key, e0, e1);
            const uint64_t mid = s & ~(ob | (1ULL << a));
            const uint32_t pout = __popcll(mid & pm_pair(a, b)) & 1u;
            const auto mview = rank_mid32(rrow, mid);
            for (int e = e0; e < e1; ++e) {
                const unsigned code = s_code[e];
                const int c = code & 0xFF, d = code >> 8;
                const uint64_t bd = 1ULL << d;
                const uint64_t inm = bd | (1ULL << c);
                if (mid & inm) continue;
                const uint32_t neg = (pout + (uint32_t)__popcll(mid & (bd - (2ULL << c)))) & 1u;
                take((e - e0) * t.nkeys + key, mview, mid | inm, neg);
            }
        }
    }
    *dsum = make_double2(dr, di);
}

template <class Rank, int NV, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_matvec_block(int64_t D, const uint64_t *__restrict__ states, CompactDev t, Rank rank,
               const double2 *__restrict__ xi, VecPtrs<NV> y)
{
    extern __shared__ uint4 lds_u4[];
    stage_image(t, lds_u4);
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= D) return;
    double2 acc[NV], dg;
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = make_double2(0.0, 0.0);
    row_block<Rank, NV>(states[i], (const unsigned char *)lds_u4, t, rank, xi, acc, &dg);
    const double2 *xr = xi + i * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        cmac(acc[v].x, acc[v].y, dg.x, dg.y, xr[v]);
        y.p[v][i] = acc[v];
    }
}

// k_matvec_block with the partner lookup of RankNarrow (J_lo staged in LDS after the image):
// the same row walk and arithmetic, so the same Y bit for bit.
template <int NV, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_matvec_block_narrow(int64_t D, const uint64_t *__restrict__ states, CompactDev t,
                      RankNarrow rank, const uint16_t *__restrict__ jlo,
                      const double2 *__restrict__ xi, VecPtrs<NV> y, int chunk)
{
    extern __shared__ uint4 lds_u4[];
    stage_image(t, lds_u4);
    const uint16_t *sjlo = stage_narrow(t, rank, jlo, lds_u4);
    __syncthreads();
    rank.sjlo = sjlo;
    int64_t i = xcd_chunk_block(chunk) * BLOCK + threadIdx.x;
    if (i >= D) return;
    double2 acc[NV], dg;
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = make_double2(0.0, 0.0);
    row_block<RankNarrow, NV>(states[i], (const unsigned char *)lds_u4, t, rank, xi, acc, &dg);
    const double2 *xr = xi + i * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        cmac(acc[v].x, acc[v].y, dg.x, dg.y, xr[v]);
        y.p[v][i] = acc[v];
    }
}

// Xi[i*NV + v] = x_v[i]
template <int NV>
__global__ void __launch_bounds__(kRowBlock)
k_interleave(int64_t D, CVecPtrs<NV> x, double2 *__restrict__ xi)
{
    int64_t i = (int64_t)blockIdx.x * kRowBlock + threadIdx.x;
    if (i >= D) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) xi[i * NV + v] = x.p[v][i];
}

// ---------------------------------------------------------------- batched matvec
// Many independent problems (own basis, own term image, own vectors) in one launch: the
// workgroup -> problem table routes each 512-row block to its problem, which stages its own
// term image. Each problem's rows are computed exactly as by k_matvec_ldsu<., true, 1>.
// use_lin: 2 = rank tables with the low half staged in LDS (RankNarrow, jlo_n), 1 = rank
// tables in global memory, 0 = binary search
struct BatchMvDesc {
    int64_t D;
    const uint64_t *states;
    const double2 *x;
    double2 *y;
    CompactDev t;
    RankLin lin;
    RankSearch search;
    RankNarrow nar;
    const uint16_t *jlo_n;
    int use_lin;
    int64_t blk0;
};

__global__ void __launch_bounds__(kLdsBlock)
k_batch_matvec(const BatchMvDesc *__restrict__ descs, const int *__restrict__ blk2prob, int chunk)
{
    extern __shared__ uint4 lds_u4[];
    const int64_t blk = xcd_chunk_block(chunk);
    const BatchMvDesc &P = descs[blk2prob[blk]];
    stage_image(P.t, lds_u4);
    RankNarrow nar = P.nar;
    if (P.use_lin == 2) nar.sjlo = stage_narrow(P.t, nar, P.jlo_n, lds_u4);
    __syncthreads();
    int64_t i = (blk - P.blk0) * blockDim.x + threadIdx.x;
    if (i >= P.D) return;
    const unsigned char *base = (const unsigned char *)lds_u4;
    double2 v = P.use_lin == 2
        ? row_one<RankNarrow, true>(i, P.states[i], base, P.t, nar, P.x, 0, INT64_MAX, 1, false)
        : P.use_lin
        ? ldsu_row<RankLin, true, 1>(i, P.states[i], base, P.t, P.lin, P.x, 0, INT64_MAX, 1)
        : ldsu_row<RankSearch, true, 1>(i, P.states[i], base, P.t, P.search, P.x, 0, INT64_MAX, 1);
    P.y[i] = v;
}

// Up to kMaxGroup problems with the same basis and the same term structure (bucket offsets
// and annihilation codes; their amplitudes differ) in one descriptor: each row is walked
// once, each partner looked up once, and every problem's amplitude and x entry applied to
// it. Per problem the arithmetic and its order are row_one's, with the diagonal read from
// that problem's own table, so each y is bitwise that problem's single matvec.
constexpr int kMaxGroup = 4;
struct BatchMvGroupDesc {
    int64_t D;
    const uint64_t *states;
    CompactDev t;   // the shared structure (the first problem's image)
    RankLin lin;
    RankSearch search;
    RankNarrow nar;
    RankNarrow2 nar2;
    const uint16_t *jlo_n;
    int use_lin;   // as in BatchMvDesc; 3 = the two-level split (RankNarrow2, jlo_n)
    int64_t blk0;
    const double2 *amp_pm[kMaxGroup];   // position-major amplitudes, [p * nkeys + key]
    const double2 *diag[kMaxGroup];
    const double2 *x[kMaxGroup];
    double2 *y[kMaxGroup];
    // the group's x vectors interleaved, xi[j * G + g] = x[g][j] (nullptr: read x[g]); a
    // partner's G entries then share one 16 G-byte run instead of G cache lines
    const double2 *xi;
};

// off_code: the codes' offset in the LDS image (the image is staged without its diagonal
// table, which the walk reads from each problem's own global table)
template <class Rank, int G, bool Inter>
__device__ __forceinline__ void row_group(int64_t i, uint64_t s, const unsigned char *base,
                                          int off_code, const BatchMvGroupDesc &P,
                                          const Rank &rank)
{
    const uint16_t *s_keyptr = (const uint16_t *)base;
    const uint16_t *s_code = (const uint16_t *)(base + off_code);
    double yr[G], yi[G], dr[G], di[G];
#pragma unroll
    for (int g = 0; g < G; ++g) yr[g] = yi[g] = dr[g] = di[g] = 0.0;
    const auto rrow = rank_row32(rank, s);
    // pmi: the entry's index in the position-major amplitude tables
    auto take = [&](int pmi, const auto &mview, uint64_t target, uint32_t neg) {
        const int64_t j = rank_find(rrow, mview, target);
        if (RankMayMiss<Rank>::value && j < 0) return;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double2 a = P.amp_pm[g][pmi];
            const double2 xv = Inter ? P.xi[j * G + g] : P.x[g][j];
            cmac(yr[g], yi[g], flip_sign(a.x, neg), flip_sign(a.y, neg), xv);
        }
    };
    auto enter = [&](int key, int &e0, int &e1) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double2 dg = P.diag[g][key];
            dr[g] += dg.x;
            di[g] += dg.y;
        }
        e0 = s_keyptr[key];
        e1 = s_keyptr[key + 1];
    };
    uint64_t rem = s;
    while (rem) {
        const int b = 63 - __builtin_clzll(rem);
        rem ^= 1ULL << b;
        const uint64_t ob = 1ULL << b;
        if (b < P.t.nkeys) {   // one-body entries: in = {c}
            int e0, e1;
            enter(b, e0, e1);
            const uint64_t mid = s & ~ob;
            const uint32_t pout = __popcll(mid & pm_one(b)) & 1u;
            const auto mview = rank_mid32(rrow, mid);
            for (int e = e0; e < e1; ++e) {
                const int c = s_code[e] & 0xFF;
                const uint64_t inm = 1ULL << c;
                if (mid & inm) continue;
                take((e - e0) * P.t.nkeys + b, mview, mid | inm, (pout + (uint32_t)__popcll(mid & pm_one(c))) & 1u);
            }
        }
        uint64_t low = s & (ob - 1ULL);
        const int kb = 64 + ((b * (b - 1)) >> 1);
        while (low) {
            const int a = 63 - __builtin_clzll(low);
            low ^= 1ULL << a;
            const int key = kb + a;
            if (key >= P.t.nkeys) continue;
            int e0, e1;
            enter(key, e0, e1);
            const uint64_t mid = s & ~(ob | (1ULL << a));
            const uint32_t pout = __popcll(mid & pm_pair(a, b)) & 1u;
            const auto mview = rank_mid32(rrow, mid);
            for (int e = e0; e < e1; ++e) {   // two-body entries: in = {c, d}, c < d
                const unsigned code = s_code[e];
                const int c = code & 0xFF, d = code >> 8;
                const uint64_t bd = 1ULL << d;
                const uint64_t inm = bd | (1ULL << c);
                if (mid & inm) continue;
                take((e - e0) * P.t.nkeys + key, mview, mid | inm,
                     (pout + (uint32_t)__popcll(mid & (bd - (2ULL << c)))) & 1u);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        cmac(yr[g], yi[g], dr[g], di[g], Inter ? P.xi[i * G + g] : P.x[g][i]);
        P.y[g][i] = make_double2(yr[g], yi[g]);
    }
}

template <int G, bool Inter>
__device__ __forceinline__ void group_rows(int64_t i, const unsigned char *base, int off_code,
                                           const BatchMvGroupDesc &P, const RankNarrow &nar,
                                           const RankNarrow2 &nar2)
{
    if (P.use_lin == 3) row_group<RankNarrow2, G, Inter>(i, P.states[i], base, off_code, P, nar2);
    else if (P.use_lin == 2) row_group<RankNarrow, G, Inter>(i, P.states[i], base, off_code, P, nar);
    else if (P.use_lin) row_group<RankLin, G, Inter>(i, P.states[i], base, off_code, P, P.lin);
    else row_group<RankSearch, G, Inter>(i, P.states[i], base, off_code, P, P.search);
}

// BLOCK rows per workgroup: 512 in the step's shared launch, 1,024 for groups of 2^18 rows and
// more, which have their own launch (their LDS rank tables and 32 waves per CU)
template <int G, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_batch_mv_group(const BatchMvGroupDesc *__restrict__ descs, const int *__restrict__ blk2grp,
                 int chunk)
{
    extern __shared__ uint4 lds_u4[];
    const int64_t blk = xcd_chunk_block(chunk);
    const BatchMvGroupDesc &P = descs[blk2grp[blk]];
    stage_image_nodiag(P.t, lds_u4);
    const CompactDev tl = image_nodiag(P.t);
    RankNarrow nar = P.nar;
    RankNarrow2 nar2 = P.nar2;
    if (P.use_lin == 2) nar.sjlo = stage_narrow(tl, nar, P.jlo_n, lds_u4);
    if (P.use_lin == 3) nar2.sjlo = stage_narrow(tl, nar2, P.jlo_n, lds_u4);
    __syncthreads();
    const int64_t i = (blk - P.blk0) * BLOCK + threadIdx.x;
    if (i >= P.D) return;
    const unsigned char *base = (const unsigned char *)lds_u4;
    if (P.xi) group_rows<G, true>(i, base, tl.off_code, P, nar, nar2);
    else group_rows<G, false>(i, base, tl.off_code, P, nar, nar2);
}

// ---------------------------------------------------------------- eigenvector consumers
// <bra basis | O | ket basis>: rows are bra states, partners are looked up in the ket basis
// (binary search: the bases may be different sectors); a diagonal term contributes where the
// row's own state is in the ket basis. y has the bra's length, x the ket's.
template <class Rank>
__global__ void __launch_bounds__(kRowBlock)
k_matvec_rect(int64_t Dbra, const uint64_t *__restrict__ bra, OpDev op, Rank ket,
              const double2 *__restrict__ x, double2 *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Dbra) return;
    const uint64_t s = bra[i];
    const int64_t self = ket(s);
    double yr = 0.0, yi = 0.0;
    visit_row(
        s, op, ket,
        [&](uint32_t e, int64_t j, int neg) {
            double2 a = op.tamp[e];
            double2 xv = x[j];
            double pr = a.x * xv.x - a.y * xv.y;
            double pi = a.x * xv.y + a.y * xv.x;
            if (neg) { yr -= pr; yi -= pi; } else { yr += pr; yi += pi; }
        },
        [&](uint32_t e) {
            if (self < 0) return;
            double2 a = op.tamp[e];
            double2 xv = x[self];
            yr += a.x * xv.x - a.y * xv.y;
            yi += a.x * xv.y + a.y * xv.x;
        });
    y[i] = make_double2(yr, yi);
}

constexpr int kRedBlock = 256;

// block-wide sum of (a, b) in a fixed order; valid in thread 0
__device__ __forceinline__ double2 block_sum_fixed(double a, double b, double2 *sh)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = make_double2(a, b);
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { r.x += sh[w].x; r.y += sh[w].y; }
    return r;
}

// One-body reduced density matrix entries rho_ij = <psi| c_i^dagger c_j |psi> (orbitals
// 0-based), one (pair, chunk of rows) per workgroup; partial[pair * nchunks + chunk].
// Sign: (-1)^(occupied orbitals above i + above j in the state with j emptied), the rule of
// PES_1rdm (particle_reduced_density_matrix.jl:4-33).
template <class Rank>
__global__ void __launch_bounds__(kRedBlock)
k_rdm1(int64_t D, const uint64_t *__restrict__ states, Rank rank, const double2 *__restrict__ psi,
       const int2 *__restrict__ pairs, int64_t rows_per_chunk, double2 *__restrict__ partial)
{
    __shared__ double2 sh[kRedBlock / 64];
    const int2 pr = pairs[blockIdx.y];
    const int oi = pr.x, oj = pr.y;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = r0 + rows_per_chunk < D ? r0 + rows_per_chunk : D;
    double ar = 0.0, ai = 0.0;
    for (int64_t x = r0 + threadIdx.x; x < r1; x += kRedBlock) {
        const uint64_t s = states[x];
        if (!((s >> oj) & 1ULL)) continue;
        const uint64_t mid = s & ~(1ULL << oj);
        if (oi != oj && ((mid >> oi) & 1ULL)) continue;
        int64_t xo = x;
        if (oi != oj) {
            xo = rank(mid | (1ULL << oi));
            if (xo < 0) continue;
        }
        const int neg = __popcll(mid & (pm_one(oi) ^ pm_one(oj))) & 1;
        const double2 a = psi[xo], b = psi[x];
        double pr_ = a.x * b.x + a.y * b.y;    // conj(a) * b
        double pi_ = a.x * b.y - a.y * b.x;
        if (neg) { ar -= pr_; ai -= pi_; } else { ar += pr_; ai += pi_; }
    }
    double2 t = block_sum_fixed(ar, ai, sh);
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// sum_x conj(f_a[x]) i_b[x] prod_{occupied k of state x, ascending} u[k] for pair (a, b) =
// blockIdx.y (a * g + b), rows in chunks (ED_step_inner_prod,
// manybody_quantum_geometry.jl:43-91).
__global__ void __launch_bounds__(kRedBlock)
k_step_gram(int64_t D, const uint64_t *__restrict__ states, const double2 *__restrict__ u,
            const double2 *const *__restrict__ vf, const double2 *const *__restrict__ vi, int g,
            int64_t rows_per_chunk, double2 *__restrict__ partial)
{
    __shared__ double2 sh[kRedBlock / 64];
    const int a = blockIdx.y / g, bb = blockIdx.y % g;
    const double2 *f = vf[a], *iv = vi[bb];
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = r0 + rows_per_chunk < D ? r0 + rows_per_chunk : D;
    double ar = 0.0, ai = 0.0;
    for (int64_t x = r0 + threadIdx.x; x < r1; x += kRedBlock) {
        double wr = 1.0, wi = 0.0;
        for (uint64_t s = states[x]; s; s &= s - 1) {
            double2 uk = u[__builtin_ctzll(s)];
            double tr = wr * uk.x - wi * uk.y;
            wi = wr * uk.y + wi * uk.x;
            wr = tr;
        }
        const double2 fv = f[x], iv_ = iv[x];
        double cr = fv.x * iv_.x + fv.y * iv_.y;   // conj(f) * i
        double ci = fv.x * iv_.y - fv.y * iv_.x;
        ar += cr * wr - ci * wi;
        ai += cr * wi + ci * wr;
    }
    double2 t = block_sum_fixed(ar, ai, sh);
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// <bra| O2 O1 |ket> (ED_bracket with two operators, EDCore/src/multiplication.jl:170-183,
// 214-223): each thread takes ket states, applies every term of O1 and then every term of O2
// from the left (multiplication.jl:19-40: in ⊆ state, mid = state & ~in, out ∩ mid = ∅,
// sign (-1)^popcount(mid & parity)), looks the final state up in the bra basis (binary
// search: it may lie outside any sector) and accumulates conj(bra_i) amp2 amp1 ket_j. No
// intermediate space is formed. Cost per ket state: S1 + (applicable O1 terms) * S2 tests.
__device__ __forceinline__ bool apply_left(uint64_t s, uint64_t out, uint64_t in, uint64_t par,
                                           uint64_t &t, uint32_t &neg)
{
    if ((s & in) != in) return false;
    const uint64_t mid = s & ~in;
    if (mid & out) return false;
    t = mid | out;
    neg = (uint32_t)__popcll(mid & par) & 1u;
    return true;
}

struct TermsDev {
    int64_t S;
    const double2 *amp;
    const uint64_t *out, *in, *par;
};

__global__ void __launch_bounds__(kRedBlock)
k_bracket2(int64_t Dk, const uint64_t *__restrict__ kstates, const double2 *__restrict__ ket,
           RankSearch bra_rank, const double2 *__restrict__ bra, TermsDev o2, TermsDev o1,
           int64_t rows_per_chunk, double2 *__restrict__ partial)
{
    __shared__ double2 sh[kRedBlock / 64];
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = r0 + rows_per_chunk < Dk ? r0 + rows_per_chunk : Dk;
    double ar = 0.0, ai = 0.0;
    for (int64_t j = r0 + threadIdx.x; j < r1; j += kRedBlock) {
        const uint64_t s = kstates[j];
        const double2 kv = ket[j];
        for (int64_t t1 = 0; t1 < o1.S; ++t1) {
            uint64_t m;
            uint32_t n1;
            if (!apply_left(s, o1.out[t1], o1.in[t1], o1.par[t1], m, n1)) continue;
            const double2 a1 = o1.amp[t1];
            const double a1r = flip_sign(a1.x, n1), a1i = flip_sign(a1.y, n1);
            for (int64_t t2 = 0; t2 < o2.S; ++t2) {
                uint64_t f;
                uint32_t n2;
                if (!apply_left(m, o2.out[t2], o2.in[t2], o2.par[t2], f, n2)) continue;
                const int64_t i = bra_rank(f);
                if (i < 0) continue;
                const double2 a2 = o2.amp[t2];
                const double a2r = flip_sign(a2.x, n2), a2i = flip_sign(a2.y, n2);
                const double pr = a2r * a1r - a2i * a1i, pi = a2r * a1i + a2i * a1r;
                const double2 b = bra[i];
                const double cr = b.x * pr + b.y * pi, ci = b.x * pi - b.y * pr;   // conj(b) p
                ar += cr * kv.x - ci * kv.y;
                ai += cr * kv.y + ci * kv.x;
            }
        }
    }
    double2 t = block_sum_fixed(ar, ai, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Particle reduced density matrix of a momentum block (PES_MomtBlock_rdm,
// MomentumED/src/analysis/particle_reduced_density_matrix.jl:37-88) as rdm = M M^H with
// M[a, b] = (-1)^occ(s, a) psi[s] for s ^ ph = a | b (a a part-A state, b its complement in
// the block B): every (basis state, A state) pair is one thread; a pair writes at most one
// entry and every entry has at most one writer. occ(s, a): occupied orbitals of s in the
// parity mask of a (scat_occ_number, EDCore/src/types/manybodystate_basis.jl:397-409).
__device__ __forceinline__ uint32_t occ_parity(uint64_t n, uint64_t m)
{
    uint64_t mask = 0;
    while (m) {
        const int t1 = __builtin_ctzll(m);
        m &= m - 1;
        const uint64_t hi = m ? (1ULL << __builtin_ctzll(m)) : 0ULL;
        m &= m - 1;
        mask += hi - (t1 >= 63 ? 0ULL : (2ULL << t1));
    }
    return (uint32_t)__popcll(n & mask) & 1u;
}

__global__ void __launch_bounds__(kRowBlock)
k_pes_fill(int64_t D, const uint64_t *__restrict__ states, const double2 *__restrict__ psi,
           int64_t nA, const uint64_t *__restrict__ A, uint64_t ph, RankSearch bsearch,
           double2 *__restrict__ M, unsigned long long *__restrict__ counts)
{
    // counts[0]: (state, part) pairs with the part inside the (transformed) state;
    // counts[1]: those whose complement is in B
    __shared__ unsigned long long block_counts[2];
    if (threadIdx.x < 2) block_counts[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long applicable = 0, found = 0;
    const int64_t n = D * nA, stride = (int64_t)gridDim.x * kRowBlock;
    for (int64_t t = (int64_t)blockIdx.x * kRowBlock + threadIdx.x; t < n; t += stride) {
        const int64_t x = t / nA, a = t - x * nA;
        const uint64_t s = states[x], sph = s ^ ph, am = A[a];
        if ((sph & am) != am) continue;
        ++applicable;
        const int64_t jb = bsearch(sph & ~am);
        if (jb < 0) continue;
        ++found;
        const double2 v = psi[x];
        const uint32_t neg = occ_parity(s, am);
        M[a * bsearch.D + jb] = make_double2(flip_sign(v.x, neg), flip_sign(v.y, neg));
    }
    if (applicable) atomicAdd(&block_counts[0], applicable);
    if (found) atomicAdd(&block_counts[1], found);
    __syncthreads();
    if (threadIdx.x < 2 && block_counts[threadIdx.x])
        atomicAdd(&counts[threadIdx.x], block_counts[threadIdx.x]);
}

// out[p] = sum over chunks of partial[p * nchunks + c], in chunk order
__global__ void __launch_bounds__(kRedBlock)
k_sum_chunks(const double2 *__restrict__ partial, int64_t nchunks, double2 *__restrict__ out)
{
    __shared__ double2 sh[kRedBlock / 64];
    const double2 *p = partial + (int64_t)blockIdx.x * nchunks;
    double ar = 0.0, ai = 0.0;
    for (int64_t c = threadIdx.x; c < nchunks; c += kRedBlock) { ar += p[c].x; ai += p[c].y; }
    double2 t = block_sum_fixed(ar, ai, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kRowBlock - 1) / kRowBlock); }

OpDev opdev(const edm_op *op, const OpTable &t)
{
    return OpDev{t.d_keyptr, t.d_in, t.d_par, t.d_amp, t.d_term, op->has_onebody ? 1 : 0};
}

bool use_lin(const edm_op *op, const edm_basis *b)
{
    if (b->rank_mode != 1) return false;
    return b->rank_derived ? op_closed(op, b) : op_conserves(op, b);
}

RankLin make_lin(const edm_basis *b)
{
    return RankLin{b->d_jlo, b->d_jhi, b->nlo, (1ULL << b->nlo) - 1ULL, b->d_jlo16};
}

RankSearch make_search(const edm_basis *b)
{
    return RankSearch{b->d_states, b->D, b->d_start, b->start_shift, b->nbuckets};
}

}  // namespace

// Every full-list term keeps the sector: same particle number in each conserved component
// and the same total momentum mod Gk on both sides. Then every target of a sector state is
// in the sector, and a complete sector's rank tables give its exact index.
static bool op_conserves_uncached(const edm_op *op, const edm_basis *b);

bool op_conserves(const edm_op *op, const edm_basis *b)
{
    for (auto &kv : op->conserve_cache)
        if (kv.first == b->id) return kv.second;
    bool r = op_conserves_uncached(op, b);
    op->conserve_cache.emplace_back(b->id, r);
    return r;
}

static bool op_closed_uncached(const edm_op *op, const edm_basis *b);

bool op_closed(const edm_op *op, const edm_basis *b)
{
    for (auto &kv : op->conserve_cache)
        if (kv.first == b->id) return kv.second;
    bool r = op_closed_uncached(op, b);
    op->conserve_cache.emplace_back(b->id, r);
    return r;
}

// One- and two-body number-conserving operators only (the bucketed tables); the forward
// table and, for a list that is not upper-Hermitian, the adjoint one.
static bool op_closed_uncached(const edm_op *op, const edm_basis *b)
{
    if (!op->fast || op->nbits != b->nbits) return false;
    if (b->D == 0) return true;
    edm_ctx *ctx = b->ctx;
    unsigned int *d_miss = (unsigned int *)ctx->alloc(sizeof(unsigned int));
    unsigned int miss = 0;
    try {
        EDM_HIP_CHECK(hipMemsetAsync(d_miss, 0, sizeof(unsigned int), ctx->stream));
        hipLaunchKernelGGL(k_closure_miss, dim3(nblk(b->D)), dim3(kRowBlock), 0, ctx->stream,
                           b->D, b->d_states, opdev(op, op->fwd), make_search(b), d_miss);
        EDM_HIP_CHECK(hipGetLastError());
        if (!op->upper) {
            hipLaunchKernelGGL(k_closure_miss, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                               ctx->stream, b->D, b->d_states, opdev(op, op->adj),
                               make_search(b), d_miss);
            EDM_HIP_CHECK(hipGetLastError());
        }
        EDM_HIP_CHECK(hipMemcpyAsync(&miss, d_miss, sizeof(miss), hipMemcpyDeviceToHost,
                                     ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->free(d_miss, sizeof(unsigned int));
        throw;
    }
    ctx->free(d_miss, sizeof(unsigned int));
    return miss == 0;
}

static bool op_conserves_uncached(const edm_op *op, const edm_basis *b)
{
    if (!b->has_spec) return false;
    const SystemSpec &sp = b->spec;
    if (op->nbits != sp.nbits) return false;
    for (int64_t e = 0; e < op->S_full; ++e) {
        uint64_t o = op->out[e], i = op->in[e];
        int64_t mo[3] = {0, 0, 0}, mi[3] = {0, 0, 0};
        int co[8] = {0}, ci[8] = {0};
        for (uint64_t v = o; v; v &= v - 1) {
            int bb = __builtin_ctzll(v);
            if (bb >= sp.nbits) return false;
            co[sp.comp_of_bit(bb)]++;
            for (int d = 0; d < sp.dim; ++d) mo[d] += sp.orb_k(bb, d);
        }
        for (uint64_t v = i; v; v &= v - 1) {
            int bb = __builtin_ctzll(v);
            if (bb >= sp.nbits) return false;
            ci[sp.comp_of_bit(bb)]++;
            for (int d = 0; d < sp.dim; ++d) mi[d] += sp.orb_k(bb, d);
        }
        for (int c = 0; c < sp.ncc; ++c)
            if (co[c] != ci[c]) return false;
        for (int d = 0; d < sp.dim; ++d) {
            int64_t diff = mo[d] - mi[d];
            if (sp.gk[d] != 0 ? (diff % sp.gk[d]) != 0 : diff != 0) return false;
        }
    }
    return true;
}

void build_search_index(edm_basis *b)
{
    // buckets of about half a state each on average, at most 2^26 (256 MB) of them
    if (b->d_start || b->D < 1024 || b->D >= (int64_t)0xFFFFFFFFLL || b->nbits < 1) return;
    int nhi = 1;
    while (nhi < 26 && ((int64_t)1 << nhi) < 2 * b->D) ++nhi;
    nhi = std::min(nhi, b->nbits);
    const int shift = b->nbits - nhi;
    const uint64_t nb = 1ULL << nhi;
    edm_ctx *ctx = b->ctx;
    uint32_t *d = (uint32_t *)ctx->alloc((nb + 1) * sizeof(uint32_t));
    const unsigned grid = (unsigned)std::min<uint64_t>((nb + 1 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_bucket_starts, dim3(grid), dim3(256), 0, ctx->stream, b->D,
                       b->d_states, shift, nb, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->free(d, (nb + 1) * sizeof(uint32_t));
        EDM_HIP_CHECK(e);
    }
    b->d_start = d;
    b->start_shift = shift;
    b->nbuckets = nb;
}

bool verify_rank_tables(edm_basis *b)
{
    edm_ctx *ctx = b->ctx;
    unsigned int *d_bad = (unsigned int *)ctx->alloc(sizeof(unsigned int));
    EDM_HIP_CHECK(hipMemsetAsync(d_bad, 0, sizeof(unsigned int), ctx->stream));
    if (b->D > 0)
        hipLaunchKernelGGL(k_verify_rank<RankLin>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                           ctx->stream, b->D, b->d_states, make_lin(b), d_bad);
    EDM_HIP_CHECK(hipGetLastError());
    unsigned int bad = 0;
    EDM_HIP_CHECK(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->free(d_bad, sizeof(unsigned int));
    return bad == 0;
}

void free_narrow_tables(edm_basis *b)
{
    if (b->d_jhi_n) b->ctx->free(b->d_jhi_n, (1ULL << b->nhi_n) * sizeof(uint32_t));
    if (b->d_jlo_n16) b->ctx->free(b->d_jlo_n16, (1ULL << kNarrowLo) * sizeof(uint16_t));
    if (b->d_n2_top) b->ctx->free(b->d_n2_top, (1ULL << kNarrow2Top) * sizeof(uint32_t));
    if (b->d_n2_rows)
        b->ctx->free(b->d_n2_rows, ((size_t)b->n2_nrows << b->n2_m) * sizeof(uint32_t));
    b->d_jhi_n = nullptr;
    b->d_jlo_n16 = nullptr;
    b->nhi_n = 0;
    b->d_n2_top = nullptr;
    b->d_n2_rows = nullptr;
    b->n2_m = 0;
    b->n2_nrows = 0;
}

// The two-level form of the split (bases of kNarrowLo + kNarrow2Top + 1 .. + 12 bits), derived
// on the host from the sorted states, uploaded, and checked on every state on the device; no
// tables (and the global-table kernel) when a state's position in its block exceeds 16 bits,
// a low part would need two positions, or the rows exceed kNarrow2RowsMax bytes.
static void build_narrow2_tables(edm_basis *b)
{
    const int m = b->nbits - kNarrowLo - kNarrow2Top;
    if (b->rank_mode != 1 || m < 1 || m > 12 || b->D <= 0 || b->D > 0xFFFFFFFFLL) return;
    edm_ctx *ctx = b->ctx;
    const int64_t D = b->D;
    std::vector<uint64_t> st((size_t)D);
    EDM_HIP_CHECK(hipMemcpyAsync(st.data(), b->d_states, (size_t)D * sizeof(uint64_t),
                                 hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const uint64_t mmask = (1ULL << m) - 1ULL, lmask = (1ULL << kNarrowLo) - 1ULL;
    std::vector<uint32_t> top((size_t)1 << kNarrow2Top, 0xFFFFFFFFu);
    std::vector<uint32_t> rows;
    std::vector<uint16_t> jlo((size_t)1 << kNarrowLo, 0);
    std::vector<uint8_t> seen((size_t)1 << kNarrowLo, 0);
    uint64_t prev_top = ~0ULL, prev_h = ~0ULL;
    int64_t first = 0, nrows = 0;
    for (int64_t i = 0; i < D; ++i) {
        const uint64_t t = st[(size_t)i], h = t >> kNarrowLo, tp = h >> m;
        if (tp != prev_top) {
            if ((size_t)(nrows + 1) << m > kNarrow2RowsMax / sizeof(uint32_t)) return;
            top[(size_t)tp] = (uint32_t)nrows++;
            rows.resize((size_t)nrows << m, 0xFFFFFFFFu);
            prev_top = tp;
        }
        if (h != prev_h) {
            rows[((size_t)(nrows - 1) << m) | (size_t)(h & mmask)] = (uint32_t)i;
            first = i;
            prev_h = h;
        }
        const int64_t pos = i - first;
        const size_t l = (size_t)(t & lmask);
        if (pos > 0xFFFF || (seen[l] && jlo[l] != (uint16_t)pos)) return;
        jlo[l] = (uint16_t)pos;
        seen[l] = 1;
    }
    b->n2_m = m;
    b->n2_nrows = nrows;
    b->d_n2_top = (uint32_t *)ctx->alloc(top.size() * sizeof(uint32_t));
    b->d_n2_rows = (uint32_t *)ctx->alloc(rows.size() * sizeof(uint32_t));
    b->d_jlo_n16 = (uint16_t *)ctx->alloc(jlo.size() * sizeof(uint16_t));
    EDM_HIP_CHECK(hipMemcpyAsync(b->d_n2_top, top.data(), top.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(b->d_n2_rows, rows.data(), rows.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(b->d_jlo_n16, jlo.data(), jlo.size() * sizeof(uint16_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    unsigned int *d_bad = (unsigned int *)ctx->alloc(sizeof(unsigned int));
    unsigned int bad = 0;
    try {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, ctx->stream, d_bad, 1, 0u);
        const RankNarrow2Global chk{b->d_n2_top, b->d_n2_rows, b->d_jlo_n16, m};
        hipLaunchKernelGGL(k_verify_rank<RankNarrow2Global>, dim3(nblk(D)), dim3(kRowBlock), 0,
                           ctx->stream, D, b->d_states, chk, d_bad);
        EDM_HIP_CHECK(hipGetLastError());
        EDM_HIP_CHECK(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost,
                                     ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->free(d_bad, sizeof(unsigned int));
        throw;
    }
    ctx->free(d_bad, sizeof(unsigned int));
    if (bad != 0) free_narrow_tables(b);
}

void build_narrow_tables(edm_basis *b)
{
    free_narrow_tables(b);
    const int nhi = b->nbits - kNarrowLo;
    if (b->rank_mode == 1 && nhi > 24 && b->D >= ((int64_t)1 << 18)) {
        build_narrow2_tables(b);
        return;
    }
    // only where the tables split wider than kNarrowLo, J_hi_n stays within 64 MB and within
    // the basis' own 8 D bytes (a small sector of a wide system keeps its tables)
    if (b->rank_mode != 1 || b->nlo <= kNarrowLo || nhi < 1 || nhi > 24 || b->D <= 0 ||
        (4ULL << nhi) > 8ULL * (uint64_t)b->D)
        return;
    edm_ctx *ctx = b->ctx;
    b->nhi_n = nhi;
    b->d_jhi_n = (uint32_t *)ctx->alloc((1ULL << nhi) * sizeof(uint32_t));
    b->d_jlo_n16 = (uint16_t *)ctx->alloc((1ULL << kNarrowLo) * sizeof(uint16_t));
    const int64_t nh = (int64_t)1 << nhi, nl32 = ((int64_t)1 << kNarrowLo) / 2;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<int64_t>((nh + 255) / 256, 65536)),
                       dim3(256), 0, ctx->stream, b->d_jhi_n, nh, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nl32 + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint32_t *)b->d_jlo_n16, nl32, 0u);
    hipLaunchKernelGGL(k_narrow_hi, dim3(nblk(b->D)), dim3(256), 0, ctx->stream, b->D,
                       b->d_states, kNarrowLo, b->d_jhi_n);
    unsigned int *d_bad = (unsigned int *)ctx->alloc(sizeof(unsigned int));
    unsigned int bad = 0;
    try {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, ctx->stream, d_bad, 1, 0u);
        hipLaunchKernelGGL(k_narrow_lo, dim3(nblk(b->D)), dim3(256), 0, ctx->stream, b->D,
                           b->d_states, kNarrowLo, b->d_jhi_n, b->d_jlo_n16, d_bad);
        const RankNarrowGlobal chk{b->d_jhi_n, b->d_jlo_n16, kNarrowLo,
                                   (1ULL << kNarrowLo) - 1ULL};
        hipLaunchKernelGGL(k_verify_rank<RankNarrowGlobal>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                           ctx->stream, b->D, b->d_states, chk, d_bad);
        EDM_HIP_CHECK(hipGetLastError());
        EDM_HIP_CHECK(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost,
                                     ctx->stream));
        EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->free(d_bad, sizeof(unsigned int));
        throw;
    }
    ctx->free(d_bad, sizeof(unsigned int));
    if (bad != 0) free_narrow_tables(b);
}

// Kernel selection for the bucketed path (all variants compute the same y; the choice only
// changes speed). Default: the LDS-staged kernel with per-row rank view, one candidate at a
// time (u1) — the fastest measured on the 6x6 N=12 sector. EDM_MATVEC=u2|u4 resolves 2|4
// candidates together; hi|lo select the LDS kernel without the per-row rank view, with
// high- or low-bit-first pair order (EDM_XCD_SWIZZLE=1 then remaps workgroups per XCD);
// basic forces the global-memory table kernel; u1a the u1 row code with the amplitudes
// staged in LDS too (k_matvec_ldsa, 1024-row workgroups; u1 when they do not fit).
// EDM_MATVEC_SPLIT=n splits the u* kernels into n passes over windows of x. Without
// EDM_MATVEC, a (operator, basis) pair that went through matvec_tune runs the tuner's pick.
static int matvec_variant()
{
    const char *v = getenv("EDM_MATVEC");
    if (!v) return 11;
    if (!strcmp(v, "basic")) return 0;
    if (!strcmp(v, "lo")) return 1;
    if (!strcmp(v, "hi")) return 2;
    if (!strcmp(v, "u1")) return 11;
    if (!strcmp(v, "u1a")) return 13;
    if (!strcmp(v, "u2")) return 12;
    if (!strcmp(v, "u4")) return 14;
    if (!strcmp(v, "u1n")) return 15;
    if (!strcmp(v, "u1n512")) return 16;
    return 11;
}

// EDM_MATVEC_LANES=1|2|4|8|16 fixes the lanes per row of the default kernel (tests compare
// them; y is the same for all), over the tuner's pick and the size-based default.
static int forced_lanes()
{
    const char *v = getenv("EDM_MATVEC_LANES");
    if (!v) return 0;
    const int n = atoi(v);
    return (n == 1 || n == 2 || n == 4 || n == 8 || n == 16) ? n : 0;
}

static int xcd_swizzle_enabled()
{
    const char *v = getenv("EDM_XCD_SWIZZLE");
    return (v && v[0] == '1') ? 1 : 0;
}

static int matvec_passes()
{
    const char *v = getenv("EDM_MATVEC_SPLIT");
    int n = v ? atoi(v) : 1;
    return n < 1 ? 1 : (n > 16 ? 16 : n);
}

// workgroup order of k_matvec_ldsu: row blocks in XCD chunks of 64 as in the LDS rank-table
// kernel (6x6 N=12 sector: 84.7 -> 82.7 ms; 48-orbital two-component sector: 8.23 -> 8.06 ms;
// chunks of 32 and 128 alike); EDM_XCD_CHUNK_LDSU=n overrides (1: the dispatch order)
static int ldsu_chunk()
{
    const char *v = getenv("EDM_XCD_CHUNK_LDSU");
    if (!v) return 64;
    const int n = atoi(v);
    return n < 1 ? 1 : (n > 4096 ? 4096 : n);
}

template <class Rank, int U>
static void launch_ldsu(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                        const double2 *x, double2 *y, int block = kLdsBlock)
{
    CompactDev cd = compact_dev(ct);
    int passes = matvec_passes();
    if (U == 1 && passes == 1 && block != kLdsBlock) {
        // workgroup size chosen by edm_matvec_tune: the same per-row code, so the same y
        unsigned nb = (unsigned)((b->D + block - 1) / block);
        if (block == 256)
            hipLaunchKernelGGL((k_matvec_ldsu<Rank, true, 1, 256>), dim3(nb), dim3(256),
                               ct.lds_bytes, ctx->stream, b->D, b->d_states, cd, rank, x, y, 0,
                               INT64_MAX, 0, 1, ldsu_chunk());
        else
            hipLaunchKernelGGL((k_matvec_ldsu<Rank, true, 1, 1024>), dim3(nb), dim3(1024),
                               ct.lds_bytes, ctx->stream, b->D, b->d_states, cd, rank, x, y, 0,
                               INT64_MAX, 0, 1, ldsu_chunk());
        return;
    }
    unsigned nb = (unsigned)((b->D + kLdsBlock - 1) / kLdsBlock);
    for (int p = 0; p < passes; ++p) {
        int64_t jmin = b->D * p / passes, jmax = b->D * (p + 1) / passes;
        if (passes == 1) jmax = INT64_MAX;   // no window
        hipLaunchKernelGGL((k_matvec_ldsu<Rank, true, U>), dim3(nb), dim3(kLdsBlock), ct.lds_bytes,
                           ctx->stream, b->D, b->d_states, cd, rank, x, y, jmin, jmax, p > 0 ? 1 : 0,
                           p == 0 ? 1 : 0, ldsu_chunk());
    }
}

template <class Rank, int BLOCK>
static void launch_ldsa_block(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                              const double2 *x, double2 *y)
{
    // LDS beyond 64 KiB needs the per-kernel opt-in, once per device (every entry point runs
    // with the context's device current)
    static std::atomic<uint64_t> attr_set{0};
    const uint64_t bit = 1ULL << (ctx->device & 63);
    if (!(attr_set.load() & bit)) {
        EDM_HIP_CHECK(hipFuncSetAttribute((const void *)k_matvec_ldsa<Rank, BLOCK>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)kLdsMaxBytes));
        attr_set.fetch_or(bit);
    }
    unsigned nb = (unsigned)((b->D + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL((k_matvec_ldsa<Rank, BLOCK>), dim3(nb), dim3(BLOCK), ct.lds_full,
                       ctx->stream, b->D, b->d_states, compact_dev(ct), rank, x, y);
}

template <class Rank>
static void launch_ldsa(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                        const double2 *x, double2 *y, int block)
{
    if (block == 256) launch_ldsa_block<Rank, 256>(ctx, b, ct, rank, x, y);
    else if (block == 512) launch_ldsa_block<Rank, 512>(ctx, b, ct, rank, x, y);
    else launch_ldsa_block<Rank, 1024>(ctx, b, ct, rank, x, y);
}

template <class Rank, int G>
static void launch_split_g(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                           const double2 *x, double2 *y)
{
    const size_t lds = (((size_t)ct.lds_bytes + 15) & ~(size_t)15) +
                       (size_t)kLdsBlock * 2 * sizeof(double2);
    if (lds > 64 * 1024) {   // per-kernel opt-in beyond 64 KiB, once per device
        static std::atomic<uint64_t> attr_set{0};
        const uint64_t bit = 1ULL << (ctx->device & 63);
        if (!(attr_set.load() & bit)) {
            EDM_HIP_CHECK(hipFuncSetAttribute((const void *)k_matvec_split<Rank, G>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)kLdsMaxBytes));
            attr_set.fetch_or(bit);
        }
    }
    const int64_t nb = (b->D * G + kLdsBlock - 1) / kLdsBlock;
    hipLaunchKernelGGL((k_matvec_split<Rank, G>), dim3((unsigned)nb), dim3(kLdsBlock), lds,
                       ctx->stream, b->D, b->d_states, compact_dev(ct), rank, x, y);
}

template <class Rank>
static void launch_split(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                         const double2 *x, double2 *y, int lanes)
{
    switch (lanes) {
    case 2: launch_split_g<Rank, 2>(ctx, b, ct, rank, x, y); break;
    case 4: launch_split_g<Rank, 4>(ctx, b, ct, rank, x, y); break;
    case 8: launch_split_g<Rank, 8>(ctx, b, ct, rank, x, y); break;
    default: launch_split_g<Rank, 16>(ctx, b, ct, rank, x, y); break;
    }
}

// Lanes per row of the default kernel when the pair was not tuned: a row per lane once there
// are 2^18 rows (half of what the device holds at once), otherwise the fewest lanes (a power
// of two, at most 4) that bring the threads to 2^18. Measured: 174 k rows run fastest with
// 2 lanes (4 and more are slower), rows of 1 k - 20 k with 4 (8 and 16 within 5 %). The
// choice only changes speed: y is bitwise the same.
static int default_kernel_code(int64_t D, bool narrow)
{
    if (narrow && D >= ((int64_t)1 << 18)) return 1024 | EDM_TUNE_NARROW;
    int lanes = 1;
    while (lanes < 4 && D * lanes < ((int64_t)1 << 18)) lanes *= 2;
    return lanes == 1 ? kLdsBlock : (kLdsBlock | (lanes << EDM_TUNE_SPLIT_SHIFT));
}

// The RankNarrow view of b's rank tables: the kNarrowLo split when it was built, or the
// tables themselves when their low half is at most kNarrowLo bits wide and has a 16-bit copy.
static bool narrow_view(const edm_basis *b, RankNarrow &r, const uint16_t *&jlo)
{
    if (b->rank_mode != 1) return false;
    if (b->d_jlo_n16 && b->d_jhi_n) {
        r = RankNarrow{b->d_jhi_n, nullptr, kNarrowLo, (1u << kNarrowLo) - 1u};
        jlo = b->d_jlo_n16;
        return true;
    }
    if (b->d_jlo16 && b->d_jhi && b->nlo >= 3 && b->nlo <= kNarrowLo && b->nhi <= 32) {
        r = RankNarrow{b->d_jhi, nullptr, b->nlo, (1u << b->nlo) - 1u};
        jlo = b->d_jlo16;
        return true;
    }
    return false;
}

// The two-level view (edm_basis::d_n2_top), when it was built.
static bool narrow2_view(const edm_basis *b, RankNarrow2 &r, const uint16_t *&jlo)
{
    if (b->rank_mode != 1 || !b->d_n2_top || !b->d_n2_rows || !b->d_jlo_n16) return false;
    r = RankNarrow2{b->d_n2_top, b->d_n2_rows, nullptr, b->n2_m, (1u << kNarrowLo) - 1u};
    jlo = b->d_jlo_n16;
    return true;
}

template <class R>
static size_t narrow_lds_bytes(const CompactTable &ct, const R &r)
{
    return (((size_t)ct.lds_bytes + 15) & ~(size_t)15) + ((size_t)r.lomask + 1) * sizeof(uint16_t);
}

// whether the LDS low-table kernel can run the pair (rank tables usable for the operator, a
// narrow view, and image + J_lo within the LDS of a CU)
static bool narrow_ok(const edm_op *op, const edm_basis *b, const CompactTable &ct)
{
    RankNarrow r;
    const uint16_t *jlo;
    return ct.ok && use_lin(op, b) && narrow_view(b, r, jlo) &&
           narrow_lds_bytes(ct, r) <= kLdsMaxBytes;
}

// the same with the two-level view (single-vector matvec only)
static bool narrow2_ok(const edm_op *op, const edm_basis *b, const CompactTable &ct)
{
    RankNarrow2 r;
    const uint16_t *jlo;
    return ct.ok && use_lin(op, b) && narrow2_view(b, r, jlo) &&
           narrow_lds_bytes(ct, r) <= kLdsMaxBytes;
}

// the single-vector LDS rank-table kernel, with either view
static bool narrow_any_ok(const edm_op *op, const edm_basis *b, const CompactTable &ct)
{
    return narrow_ok(op, b, ct) || narrow2_ok(op, b, ct);
}

// Workgroup order of the LDS rank-table kernel: each XCD takes kNarrowChunk consecutive row
// blocks of every 8 kNarrowChunk, so the blocks resident on one XCD (32 CUs x 2 workgroups)
// are neighbours and their partners share that XCD's L2, while all XCDs stay in one stretch of
// rows. 6x6 N=12 sector: 60.3-61.2 ms in dispatch order, 56.5-57.3 ms with chunks of 32 to
// 128 (y the same: only the order of workgroups changes). EDM_XCD_CHUNK=n overrides (1: the
// dispatch order).
constexpr int kNarrowChunk = 64;
static int narrow_chunk()
{
    const char *v = getenv("EDM_XCD_CHUNK");
    if (!v) return kNarrowChunk;
    const int n = atoi(v);
    return n < 1 ? 1 : (n > 4096 ? 4096 : n);
}

// Whether the LDS rank-table kernel runs its pipelined walk (each x gather lands in an LDS
// slot while the walk goes on to the next partner; the diagonal comes from global memory, so
// the image in LDS drops its diagonal table and the 16-byte slots take its place): when it
// keeps at least the workgroups per CU of the plain walk (LDS and the CU's 32 wavefronts).
// 6x6 N=12 sector: 55.3 ms against 56.3 (y the same bit for bit). EDM_NARROW_PIPE=0 turns it
// off, =2 turns it on wherever it fits the LDS.
template <class R>
static bool narrow_pipe(const CompactTable &ct, const R &r, int block)
{
    const char *v = getenv("EDM_NARROW_PIPE");
    const int mode = v ? atoi(v) : 1;
    if (mode == 0) return false;
    const size_t jlo = ((size_t)r.lomask + 1) * sizeof(uint16_t);
    const size_t plain = (((size_t)ct.lds_bytes + 15) & ~(size_t)15) + jlo;
    const CompactDev cd = image_nodiag(compact_dev(ct));
    const size_t pipe = (((size_t)cd.lds_bytes + 15) & ~(size_t)15) + jlo +
                        (size_t)block * sizeof(double2);
    if (pipe > kLdsMaxBytes) return false;
    if (mode == 2) return true;
    const size_t waves = (size_t)kCuThreads / (size_t)block;   // workgroups per CU by threads
    return std::min(kLdsMaxBytes / pipe, waves) >= std::min(kLdsMaxBytes / plain, waves);
}

template <int BLOCK, bool PIPE, class R>
static void launch_narrow_block(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct,
                                const R &r, const uint16_t *jlo, const double2 *x, double2 *y)
{
    const CompactDev cd = PIPE ? image_nodiag(compact_dev(ct)) : compact_dev(ct);
    const size_t lds = (((size_t)cd.lds_bytes + 15) & ~(size_t)15) +
                       ((size_t)r.lomask + 1) * sizeof(uint16_t);
    const size_t slots = PIPE ? (size_t)BLOCK * sizeof(double2) : 0;   // static LDS
    EDM_REQUIRE(lds + slots <= kLdsMaxBytes, EDM_EINVAL,
                "term image and rank table exceed the LDS");
    if (lds + slots > 64 * 1024) {   // per-kernel opt-in beyond 64 KiB, once per device
        static std::atomic<uint64_t> attr_set{0};
        const uint64_t bit = 1ULL << (ctx->device & 63);
        if (!(attr_set.load() & bit)) {
            EDM_HIP_CHECK(hipFuncSetAttribute((const void *)k_matvec_narrow<BLOCK, PIPE, R>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)(kLdsMaxBytes - slots)));
            attr_set.fetch_or(bit);
        }
    }
    const unsigned nb = (unsigned)((b->D + BLOCK - 1) / BLOCK);
    const int chunk = narrow_chunk();
    hipLaunchKernelGGL((k_matvec_narrow<BLOCK, PIPE, R>), dim3(nb), dim3(BLOCK), lds, ctx->stream,
                       b->D, b->d_states, compact_dev(ct), r, jlo, x, y, chunk);
}

template <class R>
static void launch_narrow_view(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct,
                               const R &r, const uint16_t *jlo, const double2 *x, double2 *y,
                               int block)
{
    const bool pipe = narrow_pipe(ct, r, block == 512 ? 512 : 1024);
    if (block == 512) {
        if (pipe) launch_narrow_block<512, true>(ctx, b, ct, r, jlo, x, y);
        else launch_narrow_block<512, false>(ctx, b, ct, r, jlo, x, y);
    } else {
        if (pipe) launch_narrow_block<1024, true>(ctx, b, ct, r, jlo, x, y);
        else launch_narrow_block<1024, false>(ctx, b, ct, r, jlo, x, y);
    }
}

static void launch_narrow(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct,
                          const double2 *x, double2 *y, int block)
{
    RankNarrow r;
    RankNarrow2 r2;
    const uint16_t *jlo = nullptr;
    if (narrow_view(b, r, jlo)) launch_narrow_view(ctx, b, ct, r, jlo, x, y, block);
    else if (narrow2_view(b, r2, jlo)) launch_narrow_view(ctx, b, ct, r2, jlo, x, y, block);
    else EDM_REQUIRE(false, EDM_EINVAL, "basis has no LDS rank split");
}

template <class Rank>
static void launch_compact(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                           const double2 *x, double2 *y, int variant, int block = kLdsBlock)
{
    // block: rows per workgroup, | EDM_TUNE_AMP_LDS for the kernel with LDS amplitudes,
    // | lanes << EDM_TUNE_SPLIT_SHIFT for the split-row kernel
    if (variant == 11 && (block >> EDM_TUNE_SPLIT_SHIFT) > 1) {
        launch_split(ctx, b, ct, rank, x, y, (block >> EDM_TUNE_SPLIT_SHIFT) & 0xFF);
        return;
    }
    if (variant == 13 && ct.lds_full > 0) {
        launch_ldsa(ctx, b, ct, rank, x, y, 1024);
        return;
    }
    if (variant == 11 && (block & EDM_TUNE_AMP_LDS) && ct.lds_full > 0) {
        launch_ldsa(ctx, b, ct, rank, x, y, block & 0xFFFF);
        return;
    }
    block &= 0xFFFF;
    if (variant >= 11) {
        if (variant == 12) launch_ldsu<Rank, 2>(ctx, b, ct, rank, x, y);
        else if (variant == 14) launch_ldsu<Rank, 4>(ctx, b, ct, rank, x, y);
        else launch_ldsu<Rank, 1>(ctx, b, ct, rank, x, y, block);
        return;
    }
    CompactDev cd = compact_dev(ct);
    unsigned nb = (unsigned)((b->D + kLdsBlock - 1) / kLdsBlock);
    int swz = xcd_swizzle_enabled();
    if (variant == 1)
        hipLaunchKernelGGL((k_matvec_lds<Rank, false>), dim3(nb), dim3(kLdsBlock), ct.lds_bytes,
                           ctx->stream, b->D, b->d_states, cd, rank, x, y, swz);
    else
        hipLaunchKernelGGL((k_matvec_lds<Rank, true>), dim3(nb), dim3(kLdsBlock), ct.lds_bytes,
                           ctx->stream, b->D, b->d_states, cd, rank, x, y, swz);
}

// The bucketed-path kernel code a matvec of (op, b) runs with the given table and variant:
// rows per workgroup | EDM_TUNE_NARROW (LDS rank-table kernel) | EDM_TUNE_AMP_LDS |
// lanes << EDM_TUNE_SPLIT_SHIFT; the tuner's pick (forward direction) over the size-based
// default, EDM_MATVEC_LANES over both, EDM_MATVEC=u1n / u1n512 forcing the LDS rank-table
// kernel where it can run.
static int bucketed_code(const edm_op *op, const edm_basis *b, ApplyMode mode,
                         const CompactTable &ct, int variant)
{
    const bool nok = narrow_any_ok(op, b, ct);
    if (variant == 15 || variant == 16)
        return nok ? ((variant == 16 ? 512 : 1024) | EDM_TUNE_NARROW) : kLdsBlock;
    int block = default_kernel_code(b->D, nok);
    for (auto &kv : op->tuned_block)
        if (mode == ApplyMode::Forward && kv.first == b->id) block = kv.second;
    if (const int lanes = forced_lanes())
        block = kLdsBlock | (lanes > 1 ? lanes << EDM_TUNE_SPLIT_SHIFT : 0);
    if ((block & EDM_TUNE_NARROW) && (!nok || variant != 11)) block = kLdsBlock;
    return block;
}

int matvec_kernel_code(const edm_op *op, const edm_basis *b)
{
    if (!op->fast || !op->cfwd.ok || b->D == 0) return 0;
    const int variant = matvec_variant();
    if (variant == 0) return 0;
    return bucketed_code(op, b, ApplyMode::Forward, op->cfwd, variant);
}

void launch_matvec(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const double2 *x,
                   double2 *y, ApplyMode mode)
{
    EDM_REQUIRE(op->nbits == b->nbits, EDM_EINVAL, "operator and basis have different bits");
    if (b->D == 0) return;
    if (op->fast) {
        const CompactTable &ct = (mode == ApplyMode::Adjoint && !op->upper) ? op->cadj : op->cfwd;
        int variant = matvec_variant();
        if (ct.ok && variant != 0) {
            const int block = bucketed_code(op, b, mode, ct, variant);
            if (block & EDM_TUNE_NARROW)
                launch_narrow(ctx, b, ct, x, y, block & 0xFFFF);
            else if (use_lin(op, b))
                launch_compact(ctx, b, ct, make_lin(b), x, y, variant >= 15 ? 11 : variant,
                               block);
            else
                launch_compact(ctx, b, ct, make_search(b), x, y, variant >= 15 ? 11 : variant,
                               block);
            EDM_HIP_CHECK(hipGetLastError());
            return;
        }
        OpDev od = opdev(op, mode == ApplyMode::Forward ? op->fwd : op->adj);
        if (use_lin(op, b))
            hipLaunchKernelGGL(k_matvec_fast<RankLin>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                               ctx->stream, b->D, b->d_states, od, make_lin(b), x, y);
        else
            hipLaunchKernelGGL(k_matvec_fast<RankSearch>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                               ctx->stream, b->D, b->d_states, od, make_search(b), x, y);
    } else {
        int adj = mode == ApplyMode::Adjoint ? 1 : 0;
        if (use_lin(op, b))
            hipLaunchKernelGGL(k_matvec_generic<RankLin>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                               ctx->stream, b->D, b->d_states, op->S_full, op->d_amp, op->d_out,
                               op->d_in, op->d_par, make_lin(b), x, y, adj);
        else
            hipLaunchKernelGGL(k_matvec_generic<RankSearch>, dim3(nblk(b->D)), dim3(kRowBlock),
                               0, ctx->stream, b->D, b->d_states, op->S_full, op->d_amp,
                               op->d_out, op->d_in, op->d_par, make_search(b), x, y, adj);
    }
    EDM_HIP_CHECK(hipGetLastError());
}

template <class Rank, int NV>
static void launch_block_nv(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                            const double2 *const *x, double2 *xi, double2 *const *y)
{
    CVecPtrs<NV> xp;
    VecPtrs<NV> yp;
    for (int v = 0; v < NV; ++v) {
        xp.p[v] = x[v];
        yp.p[v] = y[v];
    }
    hipLaunchKernelGGL(k_interleave<NV>, dim3(nblk(b->D)), dim3(kRowBlock), 0, ctx->stream, b->D,
                       xp, xi);
    CompactDev cd = compact_dev(ct);
    unsigned nb = (unsigned)((b->D + kLdsBlock - 1) / kLdsBlock);
    hipLaunchKernelGGL((k_matvec_block<Rank, NV, kLdsBlock>), dim3(nb), dim3(kLdsBlock),
                       ct.lds_bytes, ctx->stream, b->D, b->d_states, cd, rank, xi, yp);
}

template <class Rank>
static void launch_block_rank(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, Rank rank,
                              int nv, const double2 *const *x, double2 *xi, double2 *const *y)
{
    // groups of 8, then 4, 2 and 1; xi is reused by consecutive groups (stream-ordered)
    int v = 0;
    while (v < nv) {
        int n = nv - v >= 8 ? 8 : nv - v >= 4 ? 4 : nv - v >= 2 ? 2 : 1;
        if (n == 8) launch_block_nv<Rank, 8>(ctx, b, ct, rank, x + v, xi, y + v);
        else if (n == 4) launch_block_nv<Rank, 4>(ctx, b, ct, rank, x + v, xi, y + v);
        else if (n == 2) launch_block_nv<Rank, 2>(ctx, b, ct, rank, x + v, xi, y + v);
        else launch_block_nv<Rank, 1>(ctx, b, ct, rank, x + v, xi, y + v);
        EDM_HIP_CHECK(hipGetLastError());
        v += n;
    }
}

template <int NV, int BLOCK>
static void launch_block_narrow_nv(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct,
                                   const double2 *const *x, double2 *xi, double2 *const *y)
{
    RankNarrow r;
    const uint16_t *jlo = nullptr;
    EDM_REQUIRE(narrow_view(b, r, jlo), EDM_EINVAL, "basis has no LDS rank split");
    const size_t lds = narrow_lds_bytes(ct, r);
    EDM_REQUIRE(lds <= kLdsMaxBytes, EDM_EINVAL, "term image and rank table exceed the LDS");
    if (lds > 64 * 1024) {   // per-kernel opt-in beyond 64 KiB, once per device
        static std::atomic<uint64_t> attr_set{0};
        const uint64_t bit = 1ULL << (ctx->device & 63);
        if (!(attr_set.load() & bit)) {
            EDM_HIP_CHECK(hipFuncSetAttribute((const void *)k_matvec_block_narrow<NV, BLOCK>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)kLdsMaxBytes));
            attr_set.fetch_or(bit);
        }
    }
    CVecPtrs<NV> xp;
    VecPtrs<NV> yp;
    for (int v = 0; v < NV; ++v) {
        xp.p[v] = x[v];
        yp.p[v] = y[v];
    }
    hipLaunchKernelGGL(k_interleave<NV>, dim3(nblk(b->D)), dim3(kRowBlock), 0, ctx->stream, b->D,
                       xp, xi);
    const unsigned nb = (unsigned)((b->D + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL((k_matvec_block_narrow<NV, BLOCK>), dim3(nb), dim3(BLOCK), lds,
                       ctx->stream, b->D, b->d_states, compact_dev(ct), r, jlo, xi, yp,
                       narrow_chunk());
}

// EDM_BLOCK_NARROW=0 keeps the block matvec on the global-table lookup; =512 or =1024 runs
// the LDS rank-table block kernel with that workgroup wherever it can run. Default: that kernel
// from 2^18 rows on, with 1024 rows up to 4 vectors per
// launch, 512 for 8 (6x6 N=12 sector, ms per vector: 2 vectors 59.4 global table / 57.4 /
// 50.8; 4 vectors 47.9 / 42.7 / 41.1; 8 vectors 44.3 / 41.5 / 47.5).
static int block_narrow_setting()
{
    const char *v = getenv("EDM_BLOCK_NARROW");
    if (!v) return -1;
    const int n = atoi(v);
    return n == 0 ? 0 : (n == 1024 ? 1024 : 512);
}

static void launch_block_narrow(edm_ctx *ctx, const edm_basis *b, const CompactTable &ct, int nv,
                                const double2 *const *x, double2 *xi, double2 *const *y,
                                int setting)
{
    int v = 0;
    while (v < nv) {
        const int n = nv - v >= 8 ? 8 : nv - v >= 4 ? 4 : nv - v >= 2 ? 2 : 1;
        const bool wide = setting == 1024 || (setting < 0 && n < 8);
        if (n == 8) {
            if (wide) launch_block_narrow_nv<8, 1024>(ctx, b, ct, x + v, xi, y + v);
            else launch_block_narrow_nv<8, 512>(ctx, b, ct, x + v, xi, y + v);
        } else if (n == 4) {
            if (wide) launch_block_narrow_nv<4, 1024>(ctx, b, ct, x + v, xi, y + v);
            else launch_block_narrow_nv<4, 512>(ctx, b, ct, x + v, xi, y + v);
        } else if (n == 2) {
            if (wide) launch_block_narrow_nv<2, 1024>(ctx, b, ct, x + v, xi, y + v);
            else launch_block_narrow_nv<2, 512>(ctx, b, ct, x + v, xi, y + v);
        } else {
            if (wide) launch_block_narrow_nv<1, 1024>(ctx, b, ct, x + v, xi, y + v);
            else launch_block_narrow_nv<1, 512>(ctx, b, ct, x + v, xi, y + v);
        }
        EDM_HIP_CHECK(hipGetLastError());
        v += n;
    }
}

void launch_matvec_block(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int nv,
                         const double2 *const *x, double2 *xi, double2 *const *y)
{
    EDM_REQUIRE(op->nbits == b->nbits, EDM_EINVAL, "operator and basis have different bits");
    if (b->D == 0 || nv <= 0) return;
    if (!(op->fast && op->cfwd.ok && matvec_variant() == 11)) {
        for (int v = 0; v < nv; ++v) launch_matvec(ctx, op, b, x[v], y[v], ApplyMode::Forward);
        return;
    }
    const int nb = block_narrow_setting();
    if (nb != 0 && (nb > 0 || b->D >= ((int64_t)1 << 18)) && narrow_ok(op, b, op->cfwd)) {
        launch_block_narrow(ctx, b, op->cfwd, nv, x, xi, y, nb);
        return;
    }
    if (use_lin(op, b))
        launch_block_rank(ctx, b, op->cfwd, make_lin(b), nv, x, xi, y);
    else
        launch_block_rank(ctx, b, op->cfwd, make_search(b), nv, x, xi, y);
}

// XCD chunk of the batched matvec kernels' row blocks: a problem's consecutive blocks stay on
// one XCD, so its x is cached in one L2 instead of eight (3,840-problem sweep: 491 / 508 / 516
// sector-solves/s with chunks of 1 / 16 / 64); EDM_XCD_CHUNK_BATCH=n overrides (1: dispatch
// order)
static int batch_chunk()
{
    const char *v = getenv("EDM_XCD_CHUNK_BATCH");
    if (!v) return 64;
    const int n = atoi(v);
    return n < 1 ? 1 : (n > 4096 ? 4096 : n);
}

// EDM_BATCH_NARROW=0 keeps the batched matvec's rank lookups in global memory
static bool batch_narrow_enabled()
{
    const char *v = getenv("EDM_BATCH_NARROW");
    return !(v && v[0] == '0');
}
// the batched kernels stage the low rank table only while image and table stay within a
// third of a CU's LDS: one launch stages as much as its largest problem needs, and three
// 512-row workgroups per CU (24 waves) is as far as the occupancy of the whole launch may fall
constexpr size_t kBatchLdsMax = kLdsMaxBytes / 3;

// dynamic LDS beyond 64 KiB needs the per-kernel opt-in, once per device
static void batch_lds_optin(edm_ctx *ctx, const void *kernel, std::atomic<uint64_t> &flag,
                            size_t lds_bytes)
{
    if (lds_bytes <= 64 * 1024) return;
    const uint64_t bit = 1ULL << (ctx->device & 63);
    if (flag.load() & bit) return;
    EDM_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kLdsMaxBytes));
    flag.fetch_or(bit);
}

size_t batch_mv_desc_size() { return sizeof(BatchMvDesc); }

int64_t batch_mv_blocks(int64_t D, int block) { return (D + block - 1) / block; }

bool batch_mv_desc(const edm_op *op, const edm_basis *b, const double2 *x, double2 *y,
                   int64_t blk0, void *out, size_t *lds_bytes)
{
    EDM_REQUIRE(op->nbits == b->nbits, EDM_EINVAL, "operator and basis have different bits");
    const CompactTable &ct = op->cfwd;
    if (!op->fast || !ct.ok || matvec_variant() != 11 || matvec_passes() != 1) return false;
    // a problem whose own matvec is the 1,024-row LDS rank-table kernel gets its own launch of
    // it (the batch engine's lone path; same row walk, so the same y): in the shared 512-row
    // launch its image and table would halve the occupancy of every problem
    if (matvec_kernel_code(op, b) & EDM_TUNE_NARROW &&
        narrow_any_ok(op, b, ct) && b->D >= ((int64_t)1 << 18))
        return false;
    BatchMvDesc d{};
    d.D = b->D;
    d.states = b->d_states;
    d.x = x;
    d.y = y;
    d.t = compact_dev(ct);
    d.use_lin = use_lin(op, b) ? 1 : 0;
    if (d.use_lin) d.lin = make_lin(b);
    d.search = make_search(b);
    d.blk0 = blk0;
    *lds_bytes = ct.lds_bytes;
    if (batch_narrow_enabled() && d.use_lin && narrow_ok(op, b, ct) &&
        narrow_lds_bytes(ct, d.nar) <= kBatchLdsMax && narrow_view(b, d.nar, d.jlo_n)) {
        d.use_lin = 2;
        *lds_bytes = narrow_lds_bytes(ct, d.nar);
    }
    memcpy(out, &d, sizeof(d));
    return true;
}

void launch_batch_mv(edm_ctx *ctx, const void *d_descs, const int *d_blk2prob, int64_t nblocks,
                     size_t lds_bytes)
{
    if (nblocks == 0) return;
    static std::atomic<uint64_t> optin{0};
    batch_lds_optin(ctx, (const void *)k_batch_matvec, optin, lds_bytes);
    hipLaunchKernelGGL(k_batch_matvec, dim3((unsigned)nblocks), dim3(kLdsBlock), lds_bytes,
                       ctx->stream, (const BatchMvDesc *)d_descs, d_blk2prob, batch_chunk());
    EDM_HIP_CHECK(hipGetLastError());
}

int batch_mv_max_group()
{
    const char *v = getenv("EDM_BATCH_MAX_GROUP");
    const int g = v ? atoi(v) : kMaxGroup;
    return g < 1 ? 1 : (g > kMaxGroup ? kMaxGroup : g);
}

size_t batch_mv_group_desc_size() { return sizeof(BatchMvGroupDesc); }

bool batch_mv_same_shape(const edm_op *a, const edm_op *b)
{
    const CompactTable &ta = a->cfwd, &tb = b->cfwd;
    return a->fast && b->fast && ta.ok && tb.ok && a->nbits == b->nbits &&
           ta.shape_hash == tb.shape_hash && ta.nkeys == tb.nkeys && ta.noff == tb.noff &&
           ta.lds_bytes == tb.lds_bytes && ta.shape == tb.shape;
}

bool batch_mv_group_desc(const edm_op *const *ops, const edm_basis *b, const double2 *const *x,
                         double2 *const *y, int G, int64_t blk0, void *out, size_t *lds_bytes,
                         const double2 *xi, bool big)
{
    EDM_REQUIRE(G >= 2 && G <= kMaxGroup, EDM_EINVAL, "group size out of range");
    if (matvec_variant() != 11 || matvec_passes() != 1) return false;
    const int lin = use_lin(ops[0], b) ? 1 : 0;
    for (int g = 0; g < G; ++g) {
        EDM_REQUIRE(ops[g]->nbits == b->nbits, EDM_EINVAL,
                    "operator and basis have different bits");
        if (!batch_mv_same_shape(ops[0], ops[g]) || (use_lin(ops[g], b) ? 1 : 0) != lin)
            return false;
    }
    BatchMvGroupDesc d{};
    d.D = b->D;
    d.states = b->d_states;
    d.t = compact_dev(ops[0]->cfwd);
    d.use_lin = lin;
    if (lin) d.lin = make_lin(b);
    d.search = make_search(b);
    d.blk0 = blk0;
    d.xi = xi;
    for (int g = 0; g < G; ++g) {
        d.amp_pm[g] = ops[g]->cfwd.d_amp_pm;
        d.diag[g] = ops[g]->cfwd.d_diag;
        d.x[g] = x[g];
        d.y[g] = y[g];
    }
    const CompactTable &ct = ops[0]->cfwd;
    // the LDS image leaves out the diagonal table (row_group reads each problem's own)
    const size_t img = ((size_t)image_nodiag(d.t).lds_bytes + 15) & ~(size_t)15;
    const size_t jlo_bytes = ((size_t)1 << kNarrowLo) * sizeof(uint16_t);
    // a shared launch keeps three 512-row workgroups per CU; a big group's own launch two
    // 1,024-row ones
    const size_t cap = big ? kLdsMaxBytes / 2 : kBatchLdsMax;
    *lds_bytes = img;
    if (batch_narrow_enabled() && lin && narrow_ok(ops[0], b, ct) &&
        narrow_view(b, d.nar, d.jlo_n) && img + ((size_t)d.nar.lomask + 1) * 2 <= cap) {
        d.use_lin = 2;
        *lds_bytes = img + ((size_t)d.nar.lomask + 1) * sizeof(uint16_t);
    } else if (big && batch_narrow_enabled() && lin && narrow2_ok(ops[0], b, ct) &&
               narrow2_view(b, d.nar2, d.jlo_n) && img + jlo_bytes <= cap) {
        d.use_lin = 3;
        *lds_bytes = img + jlo_bytes;
    }
    memcpy(out, &d, sizeof(d));
    return true;
}

void launch_batch_interleave(edm_ctx *ctx, int G, const double2 *const *x, int64_t D,
                             double2 *xi)
{
    EDM_REQUIRE(G >= 2 && G <= kMaxGroup, EDM_EINVAL, "group size out of range");
    if (D == 0) return;
    switch (G) {
    case 2: {
        CVecPtrs<2> p{{x[0], x[1]}};
        hipLaunchKernelGGL(k_interleave<2>, dim3(nblk(D)), dim3(kRowBlock), 0, ctx->stream, D, p,
                           xi);
        break;
    }
    case 3: {
        CVecPtrs<3> p{{x[0], x[1], x[2]}};
        hipLaunchKernelGGL(k_interleave<3>, dim3(nblk(D)), dim3(kRowBlock), 0, ctx->stream, D, p,
                           xi);
        break;
    }
    default: {
        CVecPtrs<4> p{{x[0], x[1], x[2], x[3]}};
        hipLaunchKernelGGL(k_interleave<4>, dim3(nblk(D)), dim3(kRowBlock), 0, ctx->stream, D, p,
                           xi);
        break;
    }
    }
    EDM_HIP_CHECK(hipGetLastError());
}

template <int G, int BLOCK>
static void launch_group_g(edm_ctx *ctx, const BatchMvGroupDesc *dd, const int *d_blk2grp,
                           int64_t nblocks, size_t lds_bytes)
{
    static std::atomic<uint64_t> optin{0};
    batch_lds_optin(ctx, (const void *)k_batch_mv_group<G, BLOCK>, optin, lds_bytes);
    hipLaunchKernelGGL((k_batch_mv_group<G, BLOCK>), dim3((unsigned)nblocks), dim3(BLOCK),
                       lds_bytes, ctx->stream, dd, d_blk2grp, batch_chunk());
}

void launch_batch_mv_group(edm_ctx *ctx, int G, const void *d_descs, const int *d_blk2grp,
                           int64_t nblocks, size_t lds_bytes, int block)
{
    if (nblocks == 0) return;
    EDM_REQUIRE(block == kLdsBlock || block == 1024, EDM_EINVAL, "group workgroup size");
    const auto *dd = (const BatchMvGroupDesc *)d_descs;
    if (block == 1024) {
        if (G == 2) launch_group_g<2, 1024>(ctx, dd, d_blk2grp, nblocks, lds_bytes);
        else if (G == 3) launch_group_g<3, 1024>(ctx, dd, d_blk2grp, nblocks, lds_bytes);
        else launch_group_g<4, 1024>(ctx, dd, d_blk2grp, nblocks, lds_bytes);
    } else {
        if (G == 2) launch_group_g<2, kLdsBlock>(ctx, dd, d_blk2grp, nblocks, lds_bytes);
        else if (G == 3) launch_group_g<3, kLdsBlock>(ctx, dd, d_blk2grp, nblocks, lds_bytes);
        else launch_group_g<4, kLdsBlock>(ctx, dd, d_blk2grp, nblocks, lds_bytes);
    }
    EDM_HIP_CHECK(hipGetLastError());
}

void launch_connectivity(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t row0,
                         int64_t nrows, int64_t *d_counts, const int64_t *d_rowptr,
                         int64_t *d_cols, int32_t *d_terms, int8_t *d_negs)
{
    EDM_REQUIRE(op->fast, EDM_EINVAL,
                "connectivity dump supports one- and two-body number-conserving operators");
    if (nrows == 0) return;
    OpDev od = opdev(op, op->fwd);
    bool lin = use_lin(op, b);
    if (d_counts) {
        if (lin)
            hipLaunchKernelGGL(k_conn_count<RankLin>, dim3(nblk(nrows)), dim3(kRowBlock), 0,
                               ctx->stream, row0, nrows, b->d_states, od, make_lin(b), d_counts);
        else
            hipLaunchKernelGGL(k_conn_count<RankSearch>, dim3(nblk(nrows)), dim3(kRowBlock), 0,
                               ctx->stream, row0, nrows, b->d_states, od, make_search(b),
                               d_counts);
    } else {
        if (lin)
            hipLaunchKernelGGL(k_conn_fill<RankLin>, dim3(nblk(nrows)), dim3(kRowBlock), 0,
                               ctx->stream, row0, nrows, b->d_states, od, make_lin(b), d_rowptr,
                               d_cols, d_terms, d_negs);
        else
            hipLaunchKernelGGL(k_conn_fill<RankSearch>, dim3(nblk(nrows)), dim3(kRowBlock), 0,
                               ctx->stream, row0, nrows, b->d_states, od, make_search(b),
                               d_rowptr, d_cols, d_terms, d_negs);
    }
    EDM_HIP_CHECK(hipGetLastError());
}

void launch_csr_keys(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const int64_t *d_rowptr,
                     uint64_t *d_keys)
{
    EDM_REQUIRE(op->fast, EDM_EINVAL,
                "the stored matrix supports one- and two-body number-conserving operators");
    if (b->D == 0) return;
    OpDev od = opdev(op, op->fwd);
    if (use_lin(op, b))
        hipLaunchKernelGGL(k_csr_keys<RankLin>, dim3(nblk(b->D)), dim3(kRowBlock), 0, ctx->stream,
                           b->D, b->d_states, od, make_lin(b), d_rowptr, d_keys);
    else
        hipLaunchKernelGGL(k_csr_keys<RankSearch>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                           ctx->stream, b->D, b->d_states, od, make_search(b), d_rowptr, d_keys);
    EDM_HIP_CHECK(hipGetLastError());
}

int matvec_tune(edm_ctx *ctx, const edm_op *op, const edm_basis *b, double *ms)
{
    for (int k = 0; k < EDM_TUNE_CANDIDATES; ++k) ms[k] = -1.0;
    if (!op->fast || !op->cfwd.ok || matvec_variant() != 11 || matvec_passes() != 1 || b->D == 0)
        return 0;
    size_t vb = (size_t)b->D * sizeof(double2);
    double2 *x = (double2 *)ctx->alloc(vb);
    double2 *y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = kLdsBlock;
    try {
        y = (double2 *)ctx->alloc(vb);
        blas_fill_random(ctx, x, b->D, 0x7E57ULL, 1);
        EDM_HIP_CHECK(hipEventCreate(&e0));
        EDM_HIP_CHECK(hipEventCreate(&e1));
        const int blocks[3] = {256, 512, 1024};
        const bool nok = narrow_any_ok(op, b, op->cfwd);
        double bestms = 1e300;
        for (int k = 0; k < EDM_TUNE_CANDIDATES; ++k) {
            const bool amp_lds = k >= 3 && k < 6;
            if (amp_lds && op->cfwd.lds_full == 0) continue;
            if (k >= 10 && !nok) continue;
            const int code = k < 6    ? (blocks[k % 3] | (amp_lds ? EDM_TUNE_AMP_LDS : 0))
                             : k < 10 ? (kLdsBlock | ((2 << (k - 6)) << EDM_TUNE_SPLIT_SHIFT))
                                      : ((k == 10 ? 512 : 1024) | EDM_TUNE_NARROW);
            auto run = [&]() {
                if (code & EDM_TUNE_NARROW)
                    launch_narrow(ctx, b, op->cfwd, x, y, code & 0xFFFF);
                else if (use_lin(op, b))
                    launch_compact(ctx, b, op->cfwd, make_lin(b), x, y, 11, code);
                else
                    launch_compact(ctx, b, op->cfwd, make_search(b), x, y, 11, code);
            };
            run();   // warm-up
            EDM_HIP_CHECK(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < 3; ++r) run();
            EDM_HIP_CHECK(hipEventRecord(e1, ctx->stream));
            EDM_HIP_CHECK(hipEventSynchronize(e1));
            float t_ms = 0.f;
            EDM_HIP_CHECK(hipEventElapsedTime(&t_ms, e0, e1));
            ms[k] = t_ms / 3.0;
            if (ms[k] < bestms) {
                bestms = ms[k];
                best = code;
            }
        }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->free(x, vb);
        if (y) ctx->free(y, vb);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->free(x, vb);
    ctx->free(y, vb);
    bool found = false;
    for (auto &kv : op->tuned_block)
        if (kv.first == b->id) { kv.second = best; found = true; }
    if (!found) op->tuned_block.emplace_back(b->id, best);
    return best;
}

void launch_matvec_between(edm_ctx *ctx, const edm_op *op, const edm_basis *bra,
                           const edm_basis *ket, const double2 *x, double2 *y)
{
    EDM_REQUIRE(op->nbits == bra->nbits && op->nbits == ket->nbits, EDM_EINVAL,
                "operator and bases have different bits");
    if (bra == ket) {
        launch_matvec(ctx, op, bra, x, y, ApplyMode::Forward);
        return;
    }
    EDM_REQUIRE(op->fast, EDM_EINVAL,
                "matrix elements between two bases support one- and two-body number-conserving "
                "operators");
    if (bra->D == 0) return;
    hipLaunchKernelGGL(k_matvec_rect<RankSearch>, dim3(nblk(bra->D)), dim3(kRowBlock), 0,
                       ctx->stream, bra->D, bra->d_states, opdev(op, op->fwd), make_search(ket),
                       x, y);
    EDM_HIP_CHECK(hipGetLastError());
}

static int64_t chunk_rows(int64_t D, int64_t nitems)
{
    // about 2048 workgroups in total, at least one row block per chunk
    int64_t want = std::max<int64_t>(1, 2048 / std::max<int64_t>(nitems, 1));
    int64_t rows = (D + want - 1) / want;
    return std::max<int64_t>(rows, kRedBlock);
}

void launch_rdm1(edm_ctx *ctx, const edm_basis *b, const double2 *psi, const int2 *d_pairs,
                 int npairs, bool lin, double2 *d_out)
{
    if (npairs == 0) return;
    const int64_t D = b->D;
    const int64_t rows = chunk_rows(D, npairs);
    const int64_t nch = std::max<int64_t>(1, (D + rows - 1) / rows);
    double2 *part = (double2 *)ctx->partials((size_t)nch * npairs * sizeof(double2));
    dim3 grid((unsigned)nch, (unsigned)npairs);
    if (lin)
        hipLaunchKernelGGL(k_rdm1<RankLin>, grid, dim3(kRedBlock), 0, ctx->stream, D, b->d_states,
                           make_lin(b), psi, d_pairs, rows, part);
    else
        hipLaunchKernelGGL(k_rdm1<RankSearch>, grid, dim3(kRedBlock), 0, ctx->stream, D,
                           b->d_states, make_search(b), psi, d_pairs, rows, part);
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)npairs), dim3(kRedBlock), 0, ctx->stream, part,
                       nch, d_out);
    EDM_HIP_CHECK(hipGetLastError());
}

void launch_bracket2(edm_ctx *ctx, const edm_op *op2, const edm_op *op1, const edm_basis *bra,
                     const edm_basis *ket, const double2 *d_bra, const double2 *d_ket,
                     double2 *d_out)
{
    const int64_t D = ket->D;
    const int64_t rows = chunk_rows(D, 1);
    const int64_t nch = std::max<int64_t>(1, (D + rows - 1) / rows);
    double2 *part = (double2 *)ctx->partials((size_t)nch * sizeof(double2));
    TermsDev t2{op2->S_full, op2->d_amp, op2->d_out, op2->d_in, op2->d_par};
    TermsDev t1{op1->S_full, op1->d_amp, op1->d_out, op1->d_in, op1->d_par};
    hipLaunchKernelGGL(k_bracket2, dim3((unsigned)nch), dim3(kRedBlock), 0, ctx->stream, D,
                       ket->d_states, d_ket, make_search(bra), d_bra, t2, t1, rows, part);
    hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(kRedBlock), 0, ctx->stream, part, nch, d_out);
    EDM_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- allocation-free bracket
// The chunk partial sums of launch_bracket2_noalloc live in this file-scope array: one copy
// per device, held until the process ends, shared by every call on that device (calls that
// overlap on one device race). Host code reaches it only through the address
// hipGetSymbolAddress gives with the device current, looked up once per context (one context
// per device) when the context is created.
constexpr int64_t kBracketScratch = 4096;   // chunk partials; chunk_rows makes at most 2048
static __device__ __align__(256) double2 scratch[kBracketScratch];

double2 *noalloc_scratch_address()
{
    void *p = nullptr;
    EDM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    return (double2 *)p;
}

void launch_bracket2_noalloc(edm_ctx *ctx, const edm_op *op2, const edm_op *op1,
                             const edm_basis *bra, const edm_basis *ket, const double2 *d_bra,
                             const double2 *d_ket, double2 *d_out)
{
    double2 *part = ctx->d_noalloc;
    EDM_REQUIRE(part != nullptr, EDM_EINVAL, "no scratch address for this context");
    const int64_t D = ket->D;
    if (D == 0 || bra->D == 0) {   // an empty sum: the chunk sum of no partials is 0
        hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(kRedBlock), 0, ctx->stream, part,
                           (int64_t)0, d_out);
        EDM_HIP_CHECK(hipGetLastError());
        return;
    }
    const int64_t rows = chunk_rows(D, 1);
    const int64_t nch = std::max<int64_t>(1, (D + rows - 1) / rows);
    EDM_REQUIRE(nch <= kBracketScratch, EDM_EINVAL,
                "bracket needs more partial sums than the allocation-free scratch holds");
    TermsDev t2{op2->S_full, op2->d_amp, op2->d_out, op2->d_in, op2->d_par};
    TermsDev t1{op1->S_full, op1->d_amp, op1->d_out, op1->d_in, op1->d_par};
    // every chunk writes its partial before the sum reads it: nothing to clear
    hipLaunchKernelGGL(k_bracket2, dim3((unsigned)nch), dim3(kRedBlock), 0, ctx->stream, D,
                       ket->d_states, d_ket, make_search(bra), d_bra, t2, t1, rows, part);
    hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(kRedBlock), 0, ctx->stream, part, nch, d_out);
    EDM_HIP_CHECK(hipGetLastError());
}

void launch_pes_fill(edm_ctx *ctx, const edm_basis *psi_b, const double2 *d_psi, int64_t nA,
                     const uint64_t *d_A, uint64_t ph_mask, const edm_basis *B, double2 *d_M,
                     unsigned long long *d_counts)
{
    const int64_t n = psi_b->D * nA;
    if (n == 0) return;
    const int64_t blocks = std::min<int64_t>((n + kRowBlock - 1) / kRowBlock, 1 << 20);
    hipLaunchKernelGGL(k_pes_fill, dim3((unsigned)blocks),
                       dim3(kRowBlock), 0, ctx->stream, psi_b->D, psi_b->d_states, d_psi, nA,
                       d_A, ph_mask, make_search(B), d_M, d_counts);
    EDM_HIP_CHECK(hipGetLastError());
}

void launch_step_gram(edm_ctx *ctx, const edm_basis *b, const double2 *d_u,
                      const double2 *const *d_vf, const double2 *const *d_vi, int g, double2 *d_out)
{
    const int64_t D = b->D;
    const int64_t rows = chunk_rows(D, (int64_t)g * g);
    const int64_t nch = std::max<int64_t>(1, (D + rows - 1) / rows);
    double2 *part = (double2 *)ctx->partials((size_t)nch * g * g * sizeof(double2));
    hipLaunchKernelGGL(k_step_gram, dim3((unsigned)nch, (unsigned)(g * g)), dim3(kRedBlock), 0,
                       ctx->stream, D, b->d_states, d_u, d_vf, d_vi, g, rows, part);
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)(g * g)), dim3(kRedBlock), 0, ctx->stream,
                       part, nch, d_out);
    EDM_HIP_CHECK(hipGetLastError());
}

void launch_conn_totals(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t *nnz,
                        int64_t *nnz_off)
{
    EDM_REQUIRE(op->fast, EDM_EINVAL,
                "connectivity totals support one- and two-body number-conserving operators");
    unsigned long long *d = (unsigned long long *)ctx->alloc(2 * sizeof(unsigned long long));
    EDM_HIP_CHECK(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), ctx->stream));
    OpDev od = opdev(op, op->fwd);
    if (b->D > 0) {
        if (use_lin(op, b))
            hipLaunchKernelGGL(k_conn_totals<RankLin>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                               ctx->stream, b->D, b->d_states, od, make_lin(b), d);
        else
            hipLaunchKernelGGL(k_conn_totals<RankSearch>, dim3(nblk(b->D)), dim3(kRowBlock), 0,
                               ctx->stream, b->D, b->d_states, od, make_search(b), d);
        EDM_HIP_CHECK(hipGetLastError());
    }
    unsigned long long h[2] = {0, 0};
    EDM_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->free(d, 2 * sizeof(unsigned long long));
    *nnz = (int64_t)h[0];
    *nnz_off = (int64_t)h[1];
}

// ------------------------------------------------------------------ operator build
namespace {

uint64_t parity_mask(uint64_t occ)   // EDCore/src/types/scattering.jl:7-18
{
    uint64_t mask = 0;
    while (occ) {
        int tz1 = __builtin_ctzll(occ);
        occ &= occ - 1;
        int tz2 = occ ? __builtin_ctzll(occ) : 64;
        if (occ) occ &= occ - 1;
        mask += tz2 >= 64 ? 0ULL : (1ULL << tz2);
        mask -= tz1 + 1 >= 64 ? 0ULL : (1ULL << (tz1 + 1));
    }
    return mask;
}

int out_key(uint64_t outm)
{
    int pc = __builtin_popcountll(outm);
    if (pc == 1) return __builtin_ctzll(outm);
    int a = __builtin_ctzll(outm);
    int b = 63 - __builtin_clzll(outm);
    return 64 + b * (b - 1) / 2 + a;
}

void upload_table(edm_ctx *ctx, edm_op *op, OpTable &t, bool adjoint)
{
    const int64_t S = op->S_full;
    std::vector<std::vector<int64_t>> buckets(EDM_NKEYS);
    for (int64_t e = 0; e < S; ++e) {
        uint64_t outm = adjoint ? op->in[e] : op->out[e];
        buckets[out_key(outm)].push_back(e);
    }
    std::vector<uint32_t> keyptr(EDM_NKEYS + 1, 0);
    std::vector<uint64_t> tin;
    std::vector<uint64_t> tpar;
    std::vector<double2> tamp;
    std::vector<int32_t> tterm;
    tin.reserve(S); tpar.reserve(S); tamp.reserve(S); tterm.reserve(S);
    for (int k = 0; k < EDM_NKEYS; ++k) {
        keyptr[k] = (uint32_t)tin.size();
        auto &bk = buckets[k];
        // ascending annihilation mask inside a bucket
        std::stable_sort(bk.begin(), bk.end(), [&](int64_t a, int64_t b) {
            uint64_t ia = adjoint ? op->out[a] : op->in[a];
            uint64_t ib = adjoint ? op->out[b] : op->in[b];
            return ia < ib;
        });
        for (int64_t e : bk) {
            tin.push_back(adjoint ? op->out[e] : op->in[e]);
            tpar.push_back(op->par[e]);
            double im = op->amp[2 * e + 1];
            tamp.push_back(make_double2(op->amp[2 * e], adjoint ? -im : im));
            tterm.push_back((int32_t)e);
        }
    }
    keyptr[EDM_NKEYS] = (uint32_t)tin.size();
    t.nent = S;
    size_t n = std::max<size_t>(1, (size_t)S);
    t.d_keyptr = (uint32_t *)ctx->alloc(keyptr.size() * sizeof(uint32_t));
    t.d_in = (uint64_t *)ctx->alloc(n * sizeof(uint64_t));
    t.d_par = (uint64_t *)ctx->alloc(n * sizeof(uint64_t));
    t.d_amp = (double2 *)ctx->alloc(n * sizeof(double2));
    t.d_term = (int32_t *)ctx->alloc(n * sizeof(int32_t));
    hipStream_t st = ctx->stream;
    EDM_HIP_CHECK(hipMemcpyAsync(t.d_keyptr, keyptr.data(), keyptr.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st));
    if (S) {
        EDM_HIP_CHECK(hipMemcpyAsync(t.d_in, tin.data(), S * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        EDM_HIP_CHECK(hipMemcpyAsync(t.d_par, tpar.data(), S * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        EDM_HIP_CHECK(hipMemcpyAsync(t.d_amp, tamp.data(), S * sizeof(double2), hipMemcpyHostToDevice, st));
        EDM_HIP_CHECK(hipMemcpyAsync(t.d_term, tterm.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    EDM_HIP_CHECK(hipStreamSynchronize(st));
}

// Build the LDS image of the right-apply table (adjoint: keyed by `in`, conjugated amplitudes).
void upload_compact(edm_ctx *ctx, edm_op *op, CompactTable &t, bool adjoint)
{
    t = CompactTable{};
    struct Ent { int key; uint64_t inm; uint16_t code; double re, im; };
    std::vector<Ent> ents;
    std::vector<double2> diag(EDM_NKEYS, make_double2(0.0, 0.0));
    int maxkey = -1;
    for (int64_t e = 0; e < op->S_full; ++e) {
        uint64_t outm = adjoint ? op->in[e] : op->out[e];
        uint64_t inm = adjoint ? op->out[e] : op->in[e];
        double re = op->amp[2 * e], im = adjoint ? -op->amp[2 * e + 1] : op->amp[2 * e + 1];
        int key = out_key(outm);
        maxkey = std::max(maxkey, key);
        if (inm == outm) {
            diag[key].x += re;
            diag[key].y += im;
            continue;
        }
        int c = __builtin_ctzll(inm);
        int d = __builtin_popcountll(inm) == 2 ? 63 - __builtin_clzll(inm) : 0xFF;
        ents.push_back(Ent{key, inm, (uint16_t)(c | (d << 8)), re, im});
    }
    if (ents.size() > 65535) return;   // 16-bit offsets
    int nkeys = maxkey + 1;
    if (nkeys < 1) nkeys = 1;
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        return a.key < b.key || (a.key == b.key && a.inm < b.inm);
    });
    size_t off_diag = (((size_t)(nkeys + 1) * 2) + 15) & ~(size_t)15;
    size_t off_code = off_diag + (size_t)nkeys * 16;
    size_t lds = (off_code + ents.size() * 2 + 15) & ~(size_t)15;
    if (lds > 64 * 1024) return;
    // amplitude tail: distinct amplitudes, an adjoint entry pointing at its partner's value
    // with the conjugation flag (exact: only the sign bit of the imaginary part differs)
    std::vector<uint16_t> aidx(ents.size());
    std::vector<double2> uniq;
    {
        std::unordered_map<std::pair<uint64_t, uint64_t>, uint32_t, PairHash> seen;
        for (size_t q = 0; q < ents.size(); ++q) {
            uint64_t rb, ib, nb;
            const double nim = -ents[q].im;
            std::memcpy(&rb, &ents[q].re, 8);
            std::memcpy(&ib, &ents[q].im, 8);
            std::memcpy(&nb, &nim, 8);
            auto it = seen.find({rb, ib});
            if (it != seen.end()) { aidx[q] = (uint16_t)it->second; continue; }
            it = seen.find({rb, nb});
            if (it != seen.end()) { aidx[q] = (uint16_t)(it->second | 0x8000u); continue; }
            if (uniq.size() >= 0x8000) { uniq.clear(); break; }
            seen[{rb, ib}] = (uint32_t)uniq.size();
            aidx[q] = (uint16_t)uniq.size();
            uniq.push_back(make_double2(ents[q].re, ents[q].im));
        }
    }
    size_t off_ampidx = lds;
    size_t off_amps = (off_ampidx + ents.size() * 2 + 15) & ~(size_t)15;
    size_t lds_full = off_amps + uniq.size() * 16;
    const bool tail = !ents.empty() && !uniq.empty() && lds_full <= kLdsMaxBytes;
    const size_t img_bytes = tail ? lds_full : lds;
    std::vector<unsigned char> img(img_bytes, 0);
    if (tail) {
        std::memcpy(img.data() + off_ampidx, aidx.data(), aidx.size() * 2);
        std::memcpy(img.data() + off_amps, uniq.data(), uniq.size() * 16);
    }
    uint16_t *kp = (uint16_t *)img.data();
    double2 *dg = (double2 *)(img.data() + off_diag);
    uint16_t *cd = (uint16_t *)(img.data() + off_code);
    std::vector<double2> amp(std::max<size_t>(ents.size(), 1));
    int amp_rows = 0;
    size_t p = 0;
    for (int k = 0; k < nkeys; ++k) {
        kp[k] = (uint16_t)p;
        dg[k] = diag[k];
        const size_t p0 = p;
        while (p < ents.size() && ents[p].key == k) {
            cd[p] = ents[p].code;
            amp[p] = make_double2(ents[p].re, ents[p].im);
            ++p;
        }
        amp_rows = std::max(amp_rows, (int)(p - p0));
    }
    kp[nkeys] = (uint16_t)p;
    std::vector<double2> amp_pm((size_t)std::max(amp_rows, 1) * nkeys, make_double2(0.0, 0.0));
    for (int k = 0; k < nkeys; ++k)
        for (int q = kp[k]; q < kp[k + 1]; ++q)
            amp_pm[(size_t)(q - kp[k]) * nkeys + k] = amp[q];
    t.shape.assign(kp, kp + nkeys + 1);
    t.shape.insert(t.shape.end(), cd, cd + ents.size());
    {
        uint64_t h = 0x9E3779B97F4A7C15ULL ^ (uint64_t)t.shape.size();
        for (uint16_t v : t.shape) h = (h ^ v) * 0x100000001B3ULL;
        t.shape_hash = h;
    }
    t.nkeys = nkeys;
    t.noff = (int64_t)ents.size();
    t.off_diag = off_diag;
    t.off_code = off_code;
    t.lds_bytes = lds;
    if (tail) {
        t.off_ampidx = off_ampidx;
        t.off_amps = off_amps;
        t.lds_full = lds_full;
    }
    t.d_keyptr = (uint16_t *)ctx->alloc(img_bytes);
    t.d_diag = (double2 *)((unsigned char *)t.d_keyptr + off_diag);
    t.d_code = (uint16_t *)((unsigned char *)t.d_keyptr + off_code);
    t.d_amp = (double2 *)ctx->alloc(amp.size() * sizeof(double2));
    t.amp_rows = std::max(amp_rows, 1);
    t.d_amp_pm = (double2 *)ctx->alloc(amp_pm.size() * sizeof(double2));
    EDM_HIP_CHECK(hipMemcpyAsync(t.d_keyptr, img.data(), img_bytes, hipMemcpyHostToDevice,
                                 ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(t.d_amp, amp.data(), amp.size() * sizeof(double2),
                                 hipMemcpyHostToDevice, ctx->stream));
    EDM_HIP_CHECK(hipMemcpyAsync(t.d_amp_pm, amp_pm.data(), amp_pm.size() * sizeof(double2),
                                 hipMemcpyHostToDevice, ctx->stream));
    EDM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    t.ok = true;
}

void free_compact(edm_ctx *ctx, CompactTable &t)
{
    if (t.d_keyptr) ctx->free(t.d_keyptr, t.lds_full ? t.lds_full : t.lds_bytes);
    if (t.d_amp) ctx->free(t.d_amp, std::max<int64_t>(t.noff, 1) * sizeof(double2));
    if (t.d_amp_pm) ctx->free(t.d_amp_pm, (size_t)t.amp_rows * t.nkeys * sizeof(double2));
    t = CompactTable{};
}

void free_table(edm_ctx *ctx, OpTable &t)
{
    size_t n = std::max<size_t>(1, (size_t)t.nent);
    ctx->free(t.d_keyptr, (EDM_NKEYS + 1) * sizeof(uint32_t));
    ctx->free(t.d_in, n * sizeof(uint64_t));
    ctx->free(t.d_par, n * sizeof(uint64_t));
    ctx->free(t.d_amp, n * sizeof(double2));
    ctx->free(t.d_term, n * sizeof(int32_t));
    t = OpTable{};
}

}  // namespace

edm_op *build_op(edm_ctx *ctx, int nbits, const double *amp, const uint64_t *out,
                 const uint64_t *in, int64_t S, bool upper)
{
    EDM_REQUIRE(nbits >= 0 && nbits <= 64, EDM_EINVAL, "nbits must be 0..64");
    EDM_REQUIRE(S >= 0, EDM_EINVAL, "negative term count");
    EDM_REQUIRE(S == 0 || (amp && out && in), EDM_EINVAL, "null term array");
    uint64_t lim = nbits >= 64 ? ~0ULL : ((1ULL << nbits) - 1ULL);
    struct T { uint64_t in, out, par; double re, im; };
    std::vector<T> t;
    t.reserve(upper ? 2 * S : S);
    for (int64_t s = 0; s < S; ++s) {
        EDM_REQUIRE((out[s] & ~lim) == 0 && (in[s] & ~lim) == 0, EDM_EINVAL,
                    "term mask outside the operator's bits");
        if (upper)   // EDCore/src/types/scattering.jl:170 isupper; operator.jl:28-31 asserts it
            EDM_REQUIRE(in[s] > out[s] || (in[s] == out[s] && amp[2 * s + 1] == 0.0), EDM_EINVAL,
                        "All Scatter terms must be in the upper triangular.");
        uint64_t p = parity_mask(out[s]) ^ parity_mask(in[s]);   // scattering.jl:60
        t.push_back(T{in[s], out[s], p, amp[2 * s], amp[2 * s + 1]});
    }
    if (upper) {   // linear_map.jl:24-32: append adjoints of non-diagonal terms, sort
        for (int64_t s = 0; s < S; ++s)
            if (in[s] != out[s]) t.push_back(T{out[s], in[s], t[s].par, amp[2 * s], -amp[2 * s + 1]});
        std::stable_sort(t.begin(), t.end(), [](const T &a, const T &b) {
            return a.in < b.in || (a.in == b.in && a.out < b.out);   // scattering.jl:205-207
        });
    }
    edm_op *op = new edm_op();
    op->ctx = ctx;
    op->nbits = nbits;
    op->upper = upper;
    op->S_full = (int64_t)t.size();
    op->amp.resize(2 * t.size());
    op->out.resize(t.size());
    op->in.resize(t.size());
    op->par.resize(t.size());
    bool fast = true, one = false;
    for (size_t e = 0; e < t.size(); ++e) {
        op->amp[2 * e] = t[e].re;
        op->amp[2 * e + 1] = t[e].im;
        op->out[e] = t[e].out;
        op->in[e] = t[e].in;
        op->par[e] = t[e].par;
        int po = __builtin_popcountll(t[e].out), pi = __builtin_popcountll(t[e].in);
        if (po != pi || po < 1 || po > 2) fast = false;
        if (po == 1) one = true;
    }
    op->fast = fast;
    op->has_onebody = one;
    try {
        if (fast) {
            upload_table(ctx, op, op->fwd, false);
            upload_table(ctx, op, op->adj, true);
            upload_compact(ctx, op, op->cfwd, false);
            if (!upper) upload_compact(ctx, op, op->cadj, true);
        }
        size_t n = std::max<size_t>(1, t.size());
        op->d_amp = (double2 *)ctx->alloc(n * sizeof(double2));
        op->d_out = (uint64_t *)ctx->alloc(n * sizeof(uint64_t));
        op->d_in = (uint64_t *)ctx->alloc(n * sizeof(uint64_t));
        op->d_par = (uint64_t *)ctx->alloc(n * sizeof(uint64_t));
        if (!t.empty()) {
            hipStream_t st = ctx->stream;
            EDM_HIP_CHECK(hipMemcpyAsync(op->d_amp, op->amp.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice, st));
            EDM_HIP_CHECK(hipMemcpyAsync(op->d_out, op->out.data(), t.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
            EDM_HIP_CHECK(hipMemcpyAsync(op->d_in, op->in.data(), t.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
            EDM_HIP_CHECK(hipMemcpyAsync(op->d_par, op->par.data(), t.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
            EDM_HIP_CHECK(hipStreamSynchronize(st));
        }
    } catch (...) {
        destroy_op(op);
        throw;
    }
    return op;
}

void destroy_op(edm_op *op)
{
    if (!op) return;
    edm_ctx *ctx = op->ctx;
    if (op->fwd.d_keyptr) free_table(ctx, op->fwd);
    if (op->adj.d_keyptr) free_table(ctx, op->adj);
    free_compact(ctx, op->cfwd);
    free_compact(ctx, op->cadj);
    size_t n = std::max<size_t>(1, (size_t)op->S_full);
    ctx->free(op->d_amp, n * sizeof(double2));
    ctx->free(op->d_out, n * sizeof(uint64_t));
    ctx->free(op->d_in, n * sizeof(uint64_t));
    ctx->free(op->d_par, n * sizeof(uint64_t));
    delete op;
}
