/*
 * edm_oracle.c — CPU restatement of the reference's hot path, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (the momentumed package, libedm.so)
 * may link, load or call this file. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and there only to check or to time the CPU
 * algorithm beside the GPU one.
 *
 * Each function cites the reference file:line it restates (paths relative to the
 * MomentumED.jl repository root). Arithmetic follows the reference's order so the
 * results are bit-identical to a Julia run of the same code on the same term list:
 *   - complex products are (ar*xr - ai*xi, ar*xi + ai*xr) with no fused multiply-add
 *     (build with -ffp-contract=off), as Julia's Complex{Float64} `*` computes them;
 *   - y[i] accumulates terms in full-list order (linear_map.jl:109-115).
 *
 * Pins (what checks this oracle, see tests/test_oracle.py):
 *   parity_mask / apply       EDCore/test/test_multiplication.jl:1-13,39-61 vectors,
 *                             and a Jordan-Wigner dense operator built independently
 *                             (tests/test_oracle.py::test_jw_*).
 *   enumerate                 EDCore/test/test_colex_mbslist.jl:10-17 order, sector
 *                             sizes printed by the reference notebooks (201/200 for
 *                             3x5 Ne=5; 30709/30624/30704 for 4x6 Ne=8), brute force.
 *   matvec / connectivity     dense Jordan-Wigner Hamiltonian (textbook second
 *                             quantisation) on <= 12 orbitals; Hermiticity.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EXPORT __attribute__((visibility("default")))

static inline int ctz64(uint64_t v) { return v ? __builtin_ctzll(v) : 64; }
static inline int popc64(uint64_t v) { return __builtin_popcountll(v); }
/* Julia semantics: UInt64(1) << s is 0 for s >= 64. */
static inline uint64_t shl1(int s) { return s >= 64 ? 0ULL : (1ULL << s); }

/* EDCore/src/types/scattering.jl:7-18 — pairs set bits from the lowest up; each pair
 * (p<q) contributes bits p+1..q-1; an unpaired top bit p contributes bits p+1..63. */
ORC_EXPORT uint64_t orc_parity_mask(uint64_t occ)
{
    uint64_t mask = 0;
    while (occ) {
        int tz1 = ctz64(occ);
        occ &= occ - 1;
        int tz2 = ctz64(occ);
        if (occ) occ &= occ - 1;
        mask += shl1(tz2);
        mask -= shl1(tz1 + 1);
    }
    return mask;
}

/* EDCore/src/multiplication.jl:55-69 — right-apply `state * scat`:
 * needs out ⊆ state; diagonal terms return the state with no sign; otherwise
 * mid = state & ~out, needs mid ∩ in = ∅, result mid|in with sign from
 * popcount(mid & parity) (multiplication.jl:4-6). Returns 1 if nonzero. */
ORC_EXPORT int orc_apply_right(uint64_t state, uint64_t out, uint64_t in, uint64_t parity,
                               uint64_t *target, int *neg)
{
    if ((state & out) != out) return 0;
    if (in == out) { *target = state; *neg = 0; return 1; }
    uint64_t mid = state & ~out;
    if (mid & in) return 0;
    *target = mid | in;
    *neg = popc64(mid & parity) & 1;
    return 1;
}

/* EDCore/src/multiplication.jl:19-33 — left-apply `scat * state` (in ⊆ state). */
ORC_EXPORT int orc_apply_left(uint64_t state, uint64_t out, uint64_t in, uint64_t parity,
                              uint64_t *target, int *neg)
{
    return orc_apply_right(state, in, out, parity, target, neg);
}

/* EDCore/src/types/hilbert_subspace.jl:67-72 — searchsortedfirst + index_fit.
 * Returns the 0-based index, or -1 when the state is absent. */
ORC_EXPORT int64_t orc_search(const uint64_t *basis, int64_t D, uint64_t t)
{
    int64_t lo = 0, hi = D;            /* first index with basis[idx] >= t */
    while (lo < hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if (basis[mid] < t) lo = mid + 1; else hi = mid;
    }
    return (lo < D && basis[lo] == t) ? lo : -1;
}

typedef struct { uint64_t in, out, parity; double re, im; } orc_term;

static int cmp_term(const void *a, const void *b)
{
    const orc_term *x = (const orc_term *)a, *y = (const orc_term *)b;
    /* EDCore/src/types/scattering.jl:205-207 — order by (in, out). */
    if (x->in != y->in) return x->in < y->in ? -1 : 1;
    if (x->out != y->out) return x->out < y->out ? -1 : 1;
    return 0;
}

/* MomentumED/src/method/linear_map.jl:22-41 — full list for an upper-Hermitian
 * operator: the upper terms plus adjoint(s) = (conj Amp, in<->out, same parity,
 * scattering.jl:185-187) for every non-diagonal s, sorted by (in, out).
 * For a non-upper operator the list is copied unchanged. Output arrays need room
 * for 2*S entries. Returns the full-list length. */
ORC_EXPORT int64_t orc_full_list(int64_t S, const double *amp, const uint64_t *out,
                                 const uint64_t *in, int upper_hermitian,
                                 double *amp_f, uint64_t *out_f, uint64_t *in_f,
                                 uint64_t *par_f)
{
    orc_term *t = (orc_term *)malloc(sizeof(orc_term) * (size_t)(2 * S + 1));
    int64_t n = 0;
    for (int64_t s = 0; s < S; ++s) {
        uint64_t p = orc_parity_mask(out[s]) ^ orc_parity_mask(in[s]); /* scattering.jl:60 */
        t[n++] = (orc_term){in[s], out[s], p, amp[2 * s], amp[2 * s + 1]};
    }
    if (upper_hermitian) {
        for (int64_t s = 0; s < S; ++s) {
            if (in[s] == out[s]) continue;
            uint64_t p = orc_parity_mask(out[s]) ^ orc_parity_mask(in[s]);
            t[n++] = (orc_term){out[s], in[s], p, amp[2 * s], -amp[2 * s + 1]};
        }
        qsort(t, (size_t)n, sizeof(orc_term), cmp_term);
    }
    for (int64_t s = 0; s < n; ++s) {
        amp_f[2 * s] = t[s].re; amp_f[2 * s + 1] = t[s].im;
        out_f[s] = t[s].out; in_f[s] = t[s].in; par_f[s] = t[s].parity;
    }
    free(t);
    return n;
}

/* MomentumED/src/method/linear_map.jl:100-118 — y_i = Σ_s <i|s|j> x_j over the full
 * list, one row at a time (Hx). Rows [row0, row0+nrows) only; y is indexed from row0.
 * x, y are interleaved complex (re, im). */
ORC_EXPORT void orc_matvec_rows(int64_t D, const uint64_t *basis, int64_t S, const double *amp,
                                const uint64_t *out, const uint64_t *in, const uint64_t *par,
                                const double *x, double *y, int64_t row0, int64_t nrows,
                                int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t r = 0; r < nrows; ++r) {
        uint64_t st = basis[row0 + r];
        double yr = 0.0, yi = 0.0;
        for (int64_t s = 0; s < S; ++s) {
            uint64_t t; int neg;
            if (!orc_apply_right(st, out[s], in[s], par[s], &t, &neg)) continue;
            double ar = amp[2 * s], ai = amp[2 * s + 1];
            if (neg) { ar = -ar; ai = -ai; }
            if (ar == 0.0 && ai == 0.0) continue;        /* !iszero(amp) */
            int64_t j = orc_search(basis, D, t);
            if (j < 0) continue;                          /* index_fit false: dropped */
            double xr = x[2 * j], xi = x[2 * j + 1];
            yr += ar * xr - ai * xi;
            yi += ar * xi + ai * xr;
        }
        y[2 * r] = yr; y[2 * r + 1] = yi;
    }
}

/* Same as orc_matvec_rows for an explicit list of row indices (sampled rows). */
ORC_EXPORT void orc_matvec_rowlist(int64_t D, const uint64_t *basis, int64_t S, const double *amp,
                                   const uint64_t *out, const uint64_t *in, const uint64_t *par,
                                   const double *x, double *y, const int64_t *rows, int64_t nrows,
                                   int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t r = 0; r < nrows; ++r) {
        uint64_t st = basis[rows[r]];
        double yr = 0.0, yi = 0.0;
        for (int64_t s = 0; s < S; ++s) {
            uint64_t t; int neg;
            if (!orc_apply_right(st, out[s], in[s], par[s], &t, &neg)) continue;
            double ar = amp[2 * s], ai = amp[2 * s + 1];
            if (neg) { ar = -ar; ai = -ai; }
            if (ar == 0.0 && ai == 0.0) continue;
            int64_t j = orc_search(basis, D, t);
            if (j < 0) continue;
            double xr = x[2 * j], xi = x[2 * j + 1];
            yr += ar * xr - ai * xi;
            yi += ar * xi + ai * xr;
        }
        y[2 * r] = yr; y[2 * r + 1] = yi;
    }
}

/* Connectivity of rows [row0, row0+nrows): for each (row, term) pair that contributes
 * (linear_map.jl:110-113), record (column, full-list term index, sign bit).
 * Call with cols == NULL to get per-row counts; then with rowptr (exclusive prefix of
 * counts, length nrows+1) to fill cols/terms/negs. Entries per row are in full-list order. */
ORC_EXPORT void orc_connectivity(int64_t D, const uint64_t *basis, int64_t S, const double *amp,
                                 const uint64_t *out, const uint64_t *in, const uint64_t *par,
                                 int64_t row0, int64_t nrows, int64_t *counts,
                                 const int64_t *rowptr, int64_t *cols, int32_t *terms,
                                 int8_t *negs, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t r = 0; r < nrows; ++r) {
        uint64_t st = basis[row0 + r];
        int64_t c = 0, base = rowptr ? rowptr[r] : 0;
        for (int64_t s = 0; s < S; ++s) {
            uint64_t t; int neg;
            if (!orc_apply_right(st, out[s], in[s], par[s], &t, &neg)) continue;
            if (amp[2 * s] == 0.0 && amp[2 * s + 1] == 0.0) continue;
            int64_t j = orc_search(basis, D, t);
            if (j < 0) continue;
            if (cols) { cols[base + c] = j; terms[base + c] = (int32_t)s; negs[base + c] = (int8_t)neg; }
            ++c;
        }
        if (counts) counts[r] = c;
    }
}

/* ---------------------------------------------------------------------------------
 * Basis enumeration: MomentumED/src/preparation/momentum_decomposition.jl:101-152,188-262
 * with EDCore/src/colex_mbslist.jl:32-91 (colex iteration = ascending numeric order).
 *
 * system: dim, Nk, Nc_mix, Nc_conserve, k_list (dim x Nk column-major: k_list[d + dim*k]),
 *         Gk[dim], N_each[Nc_conserve] (>= 0).
 * Sectors are momentum_list = product of per-dim ranges with the first dim fastest
 * (momentum_decomposition.jl:209-219): 0..Gk-1, or [min*ΣN, max*ΣN] when Gk == 0.
 * counts[nsec] receives the per-sector sizes; when states != NULL, sector s's states are
 * written at states[offsets[s] ...] in ascending order (offsets = exclusive prefix of counts).
 * Returns the number of sectors (including empty ones), or -1 on bad input.
 * ------------------------------------------------------------------------------- */
typedef struct {
    int dim, nk, ncm, ncc;
    const int64_t *k_list, *gk, *n_each;
    int64_t kmin[4], krange[4];
    int64_t *counts;
    uint64_t *states;
    const int64_t *offsets;
    int64_t *fill;
} enum_ctx;

static int64_t mom_residue(int64_t k, int64_t g) /* init_parameter.jl:148-151 */
{
    if (g == 0) return k;
    int64_t r = k % g;
    return r < 0 ? r + g : r;
}

static void enum_component(enum_ctx *c, int comp, uint64_t acc, int accbits, const int64_t *accmom)
{
    int n = c->nk * c->ncm;                 /* orbitals per conserved component */
    if (comp < 0) {
        int64_t sec = 0, stride = 1;
        for (int d = 0; d < c->dim; ++d) {
            int64_t k = mom_residue(accmom[d], c->gk[d]) - c->kmin[d];
            if (k < 0 || k >= c->krange[d]) return;   /* not in momentum_list: dropped */
            sec += k * stride;
            stride *= c->krange[d];
        }
        if (c->states) c->states[c->offsets[sec] + c->fill[sec]++] = acc;
        else c->counts[sec]++;
        return;
    }
    int t = (int)c->n_each[comp];
    /* colex order over t-subsets of n bits: start at lowest t bits, next = Gosper step */
    if (t > n) return;
    uint64_t lim = (n == 64) ? 0 : (1ULL << n);
    uint64_t s = (t == 0) ? 0 : ((t == 64) ? ~0ULL : ((1ULL << t) - 1));
    for (;;) {
        int64_t mom[4];
        for (int d = 0; d < c->dim; ++d) mom[d] = accmom[d];
        for (uint64_t r = s; r; r &= r - 1) {             /* MBS_totalmomentum (:49-58) */
            int i = __builtin_ctzll(r);                   /* 0-based orbital in component */
            int ik = i % c->nk;                           /* mod1(i, Nk) - 1 */
            for (int d = 0; d < c->dim; ++d) mom[d] += c->k_list[d + c->dim * ik];
        }
        /* concat `acc * s` (manybodystate_basis.jl:81-83): acc holds the higher bits */
        enum_component(c, comp - 1, (acc << n) | s, accbits + n, mom);
        if (t == 0) break;
        /* Gosper's hack: next larger integer with the same popcount */
        uint64_t u = s & (~s + 1), v = s + u;
        if (v == 0) break;                                 /* overflow past 64 bits */
        s = v | (((v ^ s) >> 2) / u);
        if (lim && s >= lim) break;
    }
}

ORC_EXPORT int64_t orc_enumerate(int dim, int nk, int ncm, int ncc, const int64_t *k_list,
                                 const int64_t *gk, const int64_t *n_each, int64_t *counts,
                                 const int64_t *offsets, uint64_t *states)
{
    if (dim < 1 || dim > 4 || nk < 1 || ncm < 1 || ncc < 1 || (int64_t)nk * ncm * ncc > 64) return -1;
    enum_ctx c;
    memset(&c, 0, sizeof c);
    c.dim = dim; c.nk = nk; c.ncm = ncm; c.ncc = ncc;
    c.k_list = k_list; c.gk = gk; c.n_each = n_each;
    int64_t ntot = 0;
    for (int i = 0; i < ncc; ++i) { if (n_each[i] < 0) return -1; ntot += n_each[i]; }
    int64_t nsec = 1;
    for (int d = 0; d < dim; ++d) {
        if (gk[d] != 0) { c.kmin[d] = 0; c.krange[d] = gk[d]; }
        else {
            int64_t mn = k_list[d], mx = k_list[d];
            for (int k = 1; k < nk; ++k) {
                int64_t v = k_list[d + dim * k];
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
            c.kmin[d] = mn * ntot;
            c.krange[d] = mx * ntot - mn * ntot + 1;
        }
        nsec *= c.krange[d];
    }
    c.counts = counts;
    if (states) {
        c.states = states; c.offsets = offsets;
        c.fill = (int64_t *)calloc((size_t)nsec, sizeof(int64_t));
    } else {
        memset(counts, 0, sizeof(int64_t) * (size_t)nsec);
    }
    int64_t mom0[4] = {0, 0, 0, 0};
    /* recursion over conserved components, the last one first (highest bits), :260-290 */
    enum_component(&c, ncc - 1, 0, 0, mom0);
    if (c.fill) free(c.fill);
    return nsec;
}

/* Number of sectors without enumerating (size of counts[] for orc_enumerate). */
ORC_EXPORT int64_t orc_num_sectors(int dim, int nk, int ncc, const int64_t *k_list,
                                   const int64_t *gk, const int64_t *n_each)
{
    int64_t ntot = 0, nsec = 1;
    for (int i = 0; i < ncc; ++i) ntot += n_each[i];
    for (int d = 0; d < dim; ++d) {
        if (gk[d] != 0) nsec *= gk[d];
        else {
            int64_t mn = k_list[d], mx = k_list[d];
            for (int k = 1; k < nk; ++k) {
                int64_t v = k_list[d + dim * k];
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
            nsec *= mx * ntot - mn * ntot + 1;
        }
    }
    return nsec;
}

/* EDCore/src/multiplication.jl:170-183,214-223 — <bra| op2 op1 |ket> with both operators as
 * explicit (full) term lists: for every (scat2, scat1) pair and every ket state j,
 * (amp, f) = scat2 * (scat1 * state_j) (left application, multiplication.jl:19-40); f is looked
 * up in the bra basis (dropped when absent) and conj(bra[i]) * amp * ket[j] is accumulated, in
 * the reference's loop order (scat2 outer, scat1 inner, states innermost). out[0..1] = sum. */
ORC_EXPORT void orc_bracket2(int64_t Db, const uint64_t *bbasis, const double *bra,
                             int64_t Dk, const uint64_t *kbasis, const double *ket,
                             int64_t S2, const double *amp2, const uint64_t *out2,
                             const uint64_t *in2, const uint64_t *par2,
                             int64_t S1, const double *amp1, const uint64_t *out1,
                             const uint64_t *in1, const uint64_t *par1, double *out)
{
    double sr = 0.0, si = 0.0;
    for (int64_t t2 = 0; t2 < S2; ++t2)
        for (int64_t t1 = 0; t1 < S1; ++t1)
            for (int64_t j = 0; j < Dk; ++j) {
                uint64_t m, f;
                int n1, n2;
                if (!orc_apply_left(kbasis[j], out1[t1], in1[t1], par1[t1], &m, &n1)) continue;
                if (!orc_apply_left(m, out2[t2], in2[t2], par2[t2], &f, &n2)) continue;
                double a1r = n1 ? -amp1[2 * t1] : amp1[2 * t1];
                double a1i = n1 ? -amp1[2 * t1 + 1] : amp1[2 * t1 + 1];
                double a2r = n2 ? -amp2[2 * t2] : amp2[2 * t2];
                double a2i = n2 ? -amp2[2 * t2 + 1] : amp2[2 * t2 + 1];
                double ar = a2r * a1r - a2i * a1i, ai = a2r * a1i + a2i * a1r;   /* amp2 * amp1 */
                if (ar == 0.0 && ai == 0.0) continue;
                int64_t i = orc_search(bbasis, Db, f);
                if (i < 0) continue;
                double br = bra[2 * i], bi = -bra[2 * i + 1];                   /* conj(bra) */
                double kr = ket[2 * j], ki = ket[2 * j + 1];
                double cr = br * ar - bi * ai, ci = br * ai + bi * ar;
                sr += cr * kr - ci * ki;
                si += cr * ki + ci * kr;
            }
    out[0] = sr;
    out[1] = si;
}

/* EDCore/src/types/manybodystate_basis.jl:397-409 scat_occ_number: occupied orbitals of n
 * in the parity mask of i_mask (its set bits paired from the lowest: bits strictly between
 * the two of a pair; an unpaired top bit covers everything above it). */
static int occ_number(uint64_t n, uint64_t i_mask)
{
    uint64_t mask = 0;
    while (i_mask) {
        int t1 = ctz64(i_mask);
        i_mask &= i_mask - 1;
        int t2 = ctz64(i_mask);
        i_mask &= i_mask - 1;
        mask += shl1(t2);
        mask -= shl1(t1 + 1);
    }
    return popc64(n & mask);
}

/* MomentumED/src/analysis/particle_reduced_density_matrix.jl:37-88 PES_MomtBlock_rdm, in
 * the reference's loop order: for every pair (anni = A[i], crea = A[j]), i <= j, and every
 * state s of psi's basis: s_ph = s ^ ph_mask; anni must be occupied in s_ph and crea empty
 * after removing it; s_out = ((s_ph & ~anni) | crea) ^ ph_mask must be in the basis; then
 * rdm[i, j] += (-1)^(occ(s, anni) + occ(s_out, crea)) conj(psi[s_out]) psi[s]. The lower
 * triangle is filled as the conjugate of the upper one (Hermitian(rdm, :U)). rdm: nA*nA
 * complex, row-major. */
ORC_EXPORT void orc_pes_block_rdm(int64_t D, const uint64_t *basis, const double *psi,
                                  int64_t nA, const uint64_t *A, uint64_t ph_mask, double *rdm)
{
    for (int64_t i = 0; i < nA; ++i)
        for (int64_t j = 0; j < nA; ++j) {
            if (i > j) continue;
            double sr = 0.0, si = 0.0;
            const uint64_t anni = A[i], crea = A[j];
            for (int64_t x = 0; x < D; ++x) {
                uint64_t s = basis[x], s_ph = s ^ ph_mask;
                if ((s_ph & anni) != anni) continue;
                uint64_t mid = s_ph & ~anni;
                if (mid & crea) continue;
                uint64_t s_out = (mid | crea) ^ ph_mask;
                int64_t y = orc_search(basis, D, s_out);
                if (y < 0) continue;
                double pr = psi[2 * y] * psi[2 * x] + psi[2 * y + 1] * psi[2 * x + 1];
                double pi = psi[2 * y] * psi[2 * x + 1] - psi[2 * y + 1] * psi[2 * x];
                if ((occ_number(s, anni) + occ_number(s_out, crea)) & 1) { sr -= pr; si -= pi; }
                else { sr += pr; si += pi; }
            }
            rdm[2 * (i * nA + j)] = sr;
            rdm[2 * (i * nA + j) + 1] = si;
        }
    for (int64_t i = 0; i < nA; ++i)
        for (int64_t j = 0; j < i; ++j) {
            rdm[2 * (i * nA + j)] = rdm[2 * (j * nA + i)];
            rdm[2 * (i * nA + j) + 1] = -rdm[2 * (j * nA + i) + 1];
        }
}
