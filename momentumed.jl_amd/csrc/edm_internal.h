// edm_internal.h — shared internal declarations of libedm.so (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <complex>
#include <functional>
#include <string>
#include <vector>

#include "../../include/edm.h"

// ------------------------------------------------------------------ errors
struct EdmError {
    int code;
    std::string msg;
};

void edm_set_global_error(const std::string &msg);

#define EDM_HIP_CHECK(expr)                                                                 \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            throw EdmError{EDM_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)};    \
    } while (0)

#define EDM_REQUIRE(cond, code, msg)                                                        \
    do {                                                                                    \
        if (!(cond)) throw EdmError{(code), (msg)};                                         \
    } while (0)

// ------------------------------------------------------------------ system description
struct SystemSpec {
    int dim = 0, nk = 0, ncm = 1, ncc = 1;
    std::vector<int64_t> k_list;   // dim x nk, column-major
    std::vector<int64_t> gk;       // dim
    std::vector<int64_t> n_each;   // ncc
    // derived
    int nbits_comp = 0;            // orbitals per conserved component = nk*ncm
    int nbits = 0;                 // total bits = nk*ncm*ncc
    int64_t ntot = 0;
    std::vector<int64_t> kmin, krange;   // per dim sector range
    int64_t nsec = 0;

    static SystemSpec from(const edm_system *s);
    // momentum of orbital bit `b` (0-based global bit) in dimension d
    int64_t orb_k(int b, int d) const { return k_list[d + dim * ((b % nbits_comp) % nk)]; }
    int comp_of_bit(int b) const { return b / nbits_comp; }
    // sector index of a state's (unreduced) momentum sums, -1 if outside momentum_list
    int64_t sector_of_sums(const int64_t *sums) const;
    void sector_momentum(int64_t sector, int64_t *k) const;
};

// ------------------------------------------------------------------ context
struct edm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;     // 0 = unlimited
    uint64_t used = 0;
    std::string last_error;
    // binomial table C(a,b), a<=64, b<=64 on the device (enumeration)
    uint64_t *d_binom = nullptr;
    // reduction scratch
    double *d_partials = nullptr;
    size_t partials_bytes = 0;
    // this device's address of the file-scope array edm_bracket2_noalloc keeps its partials in
    double2 *d_noalloc = nullptr;

    void *alloc(size_t bytes);
    void free(void *p, size_t bytes);
    double *partials(size_t bytes);
};

// RAII set-device guard
struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != prev) (void)hipSetDevice(prev);
    }
};

// ------------------------------------------------------------------ basis
uint64_t edm_next_id();

struct edm_basis {
    uint64_t id = edm_next_id();
    edm_ctx *ctx = nullptr;
    int64_t D = 0;
    int nbits = 0;
    uint64_t *d_states = nullptr;
    // rank tables (Lin two-table): index(t) = jhi[t >> nlo] + jlo[t & ((1<<nlo)-1)]
    int rank_mode = 0;      // 0 = binary search, 1 = rank tables
    int nlo = 0, nhi = 0;
    uint32_t *d_jlo = nullptr, *d_jhi = nullptr;
    uint16_t *d_jlo16 = nullptr;   // 16-bit copy of d_jlo when every entry fits
    // the same rank split at bit kNarrowLo instead of nbits/2 (built beside rank tables whose
    // low half is wider than that): J_lo then has 2^kNarrowLo 16-bit entries, small enough to
    // be staged in LDS by the matvec, and J_hi 2^(nbits - kNarrowLo) entries
    uint32_t *d_jhi_n = nullptr;
    uint16_t *d_jlo_n16 = nullptr;
    int nhi_n = 0;
    // where J_hi over nbits - kNarrowLo bits would be too large (bases above 38 bits), the same
    // split with J_hi itself in two levels: the high part h = t >> kNarrowLo is cut into its top
    // 24 bits and n2_m low bits; d_n2_top[top] is the row of that top value (tops present in
    // the basis, in order) and d_n2_rows[row << n2_m | (h & mask)] the first index of the states
    // with high part h. J_lo is d_jlo_n16.
    uint32_t *d_n2_top = nullptr;
    uint32_t *d_n2_rows = nullptr;
    int n2_m = 0;
    int64_t n2_nrows = 0;
    bool has_spec = false;
    SystemSpec spec;
    int64_t sector = -1;
    // rank tables derived from the state list alone (no system given): valid for the states
    // themselves; an operator uses them only after a device check that every target of every
    // state is in the basis (op_closed)
    bool rank_derived = false;
    // bucket index of the binary search: start[h] = first state whose bits above start_shift
    // are >= h, for h < nbuckets, and start[nbuckets] = D (build_search_index)
    uint32_t *d_start = nullptr;
    int start_shift = 0;
    uint64_t nbuckets = 0;
};

// ------------------------------------------------------------------ operator
// Right-apply term table keyed by the creation mask (out): key a (one-body out = {a}) or
// 64 + b(b-1)/2 + a (two-body out = {a<b}). Entries of a key are in ascending `in` order.
constexpr int EDM_NKEYS = 64 + 64 * 63 / 2;   // 2080

struct OpTable {
    uint32_t *d_keyptr = nullptr;   // EDM_NKEYS + 1
    uint64_t *d_in = nullptr;       // annihilation mask per entry
    uint64_t *d_par = nullptr;      // parity mask per entry
    double2 *d_amp = nullptr;       // amplitude per entry
    int32_t *d_term = nullptr;      // index into the canonical full list
    int64_t nent = 0;
};

// Compact form of the same table for the LDS-staged kernel: off-diagonal entries only, each
// annihilation mask coded in 16 bits (c | d << 8, d = 0xFF for one-body; c < d), bucket
// offsets in 16 bits, and the diagonal amplitude of every key in its own array. Parity masks
// are recomputed from the four bit positions, so an entry costs 2 bytes of LDS plus its
// amplitude in global memory.
struct CompactTable {
    bool ok = false;
    int nkeys = 0;                  // keys >= nkeys have no entries
    int64_t noff = 0;               // off-diagonal entries
    uint16_t *d_keyptr = nullptr;   // nkeys + 1
    uint16_t *d_code = nullptr;     // noff
    double2 *d_amp = nullptr;       // noff
    double2 *d_diag = nullptr;      // nkeys (zero where a key has no diagonal term)
    // the same amplitudes position-major: entry p of key k at [p * nkeys + k], for p below the
    // longest bucket (amp_rows); rows that walk the same entry position of neighbouring keys
    // read neighbouring words
    double2 *d_amp_pm = nullptr;
    int amp_rows = 0;
    size_t lds_bytes = 0;
    size_t off_diag = 0, off_code = 0;   // byte offsets inside the LDS image
    // optional tail of the same image: a 16-bit index per entry into a table of distinct
    // amplitudes (bit 15: use the complex conjugate), so the amplitudes can be staged in LDS
    // too; lds_full = bytes of image + tail (0 when the tail does not fit in LDS)
    size_t off_ampidx = 0, off_amps = 0, lds_full = 0;
    // host copy of the table's structure (bucket offsets, then codes) and its hash: two
    // operators whose tables match here differ only in amplitudes
    std::vector<uint16_t> shape;
    uint64_t shape_hash = 0;
};

struct edm_op {
    uint64_t id = edm_next_id();
    edm_ctx *ctx = nullptr;
    // cached op_conserves() results keyed by basis id
    mutable std::vector<std::pair<uint64_t, bool>> conserve_cache;
    // matvec kernel per basis id chosen by matvec_tune: rows per workgroup (default 512),
    // | EDM_TUNE_AMP_LDS when the amplitudes are staged in LDS, | EDM_TUNE_NARROW when the
    // low-half rank table is, | lanes per row << EDM_TUNE_SPLIT_SHIFT for the split-row kernel
    mutable std::vector<std::pair<uint64_t, int>> tuned_block;
    int nbits = 0;
    bool upper = false;
    int64_t S_full = 0;
    // canonical full list (host copy, (in,out)-sorted)
    std::vector<double> amp;        // 2*S_full
    std::vector<uint64_t> out, in, par;
    bool fast = false;              // every term is 1- or 2-body and number conserving
    bool has_onebody = false;
    OpTable fwd;                    // H: right-apply keyed by out
    OpTable adj;                    // H^dagger: keyed by in (swapped roles, conj amp)
    CompactTable cfwd, cadj;        // LDS-staged forms of fwd / adj
    // generic fallback: full list SoA on the device
    double2 *d_amp = nullptr;
    uint64_t *d_out = nullptr, *d_in = nullptr, *d_par = nullptr;
};

// ------------------------------------------------------------------ batched execution
// One problem's view for the batched vector kernels (edm_blas.hip).
struct BatchVecDesc {
    int64_t D;
    const double2 *const *V;   // device pointer table of the problem's basis vectors
    int nvec;                  // orthogonalise against V[0..nvec)
    double2 *w;                // the vector being orthogonalised / normalised
    double2 *partial;          // the problem's partial-sum region, [nvec][ntiles]
    int64_t ntiles;            // reduction tiles (1024 rows each)
    int64_t tile0;             // first global tile of the problem
    int64_t blk0;              // first global 256-row block of the problem
    double2 *h, *h2;           // projection coefficients of the two Gram-Schmidt passes
    double *beta;              // norm of w
};

// One (problem, column chunk) job of the batched basis rotation (edm_blas.hip):
// out[c0 + c] = sum_i V[i] * Y[i + ldy * (c0 + c)], c < kc.
struct BatchRotDesc {
    int64_t D;
    const double2 *const *V;
    double2 *const *out;
    const double *Y;
    int m, ldy, c0, kc;
    int64_t blk0;              // first global 256-row block of the job
};

constexpr int kBatchRowBlock = 256;    // rows per block of the batched vector kernels
constexpr int kBatchTile = 1024;       // rows per reduction tile of the batched vector kernels
constexpr int kBatchRotChunk = 8;      // output columns per rotation job

// ------------------------------------------------------------------ kernels (launchers)
// enumeration (edm_enum.hip)
void enum_sector_counts(edm_ctx *ctx, const SystemSpec &sp, std::vector<int64_t> &counts);
// enumerate one sector; returns a device array of *D states (allocated on ctx)
uint64_t *enum_sector_states(edm_ctx *ctx, const SystemSpec &sp, int64_t sector, int64_t *D);
// enumerate every sector at once: one counting pass per 64 sectors and one writing pass per
// 16 sectors over the rank space; outs[k] (allocated on ctx, D[k] states) for every k < nsec
void enum_all_sector_states(edm_ctx *ctx, const SystemSpec &sp, std::vector<uint64_t *> &outs,
                            std::vector<int64_t> &D);
// rank tables (edm_rank.cpp)
bool build_rank_tables(edm_basis *b);
// the same for many complete sectors of one system (J_lo shared, J_hi on host threads);
// entry k tells whether bs[k] got tables
std::vector<bool> build_rank_tables_many(const std::vector<edm_basis *> &bs);
// rank tables from the sorted state list alone (states_host[D]), when it splits into equal
// blocks: J_hi[hi] = first index with that high half, J_lo[lo] = the position of lo in its
// block, the same in every block that holds it. Uploads them and returns true, or false.
bool derive_rank_tables(edm_basis *b, const uint64_t *states_host);
// matvec (edm_matvec.hip)
enum class ApplyMode { Forward, Adjoint };
void launch_matvec(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const double2 *x,
                   double2 *y, ApplyMode mode);
// the kernel code edm_matvec(op, b) launches (edm.h edm_matvec_kernel)
int matvec_kernel_code(const edm_op *op, const edm_basis *b);
// y_v = H x_v for v < nv. When the operator has an LDS term image, x_0..x_{nv-1} are first
// interleaved into xi (nv * D complex, caller-provided) and each group of up to 8 vectors is
// one launch of the block kernel; otherwise nv single matvecs (xi unused). Column v equals
// launch_matvec(x_v) bit for bit either way.
void launch_matvec_block(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int nv,
                         const double2 *const *x, double2 *xi, double2 *const *y);
void launch_connectivity(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t row0,
                         int64_t nrows, int64_t *d_counts, const int64_t *d_rowptr,
                         int64_t *d_cols, int32_t *d_terms, int8_t *d_negs);
bool op_conserves(const edm_op *op, const edm_basis *b);
// every target of every basis state under the operator's full list, in both directions, is
// in the basis (checked on the device by binary search, cached per basis)
bool op_closed(const edm_op *op, const edm_basis *b);
void launch_conn_totals(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t *nnz,
                        int64_t *nnz_off);
edm_op *build_op(edm_ctx *ctx, int nbits, const double *amp, const uint64_t *out,
                 const uint64_t *in, int64_t S, bool upper);
void destroy_op(edm_op *op);
bool verify_rank_tables(edm_basis *b);
// rank split of the LDS low-table matvec: J_lo over the lowest kNarrowLo bits
constexpr int kNarrowLo = 14;
// For a basis with verified rank tables whose low half is wider than kNarrowLo bits: derive
// the kNarrowLo split from the state list on the device (J_hi_n[h] = first state with high
// part h, J_lo_n[l] = position of l in its block), check it on every state and keep it, or
// free it when a check fails. High parts wider than 24 bits (J_hi_n over 64 MB) get none.
void build_narrow_tables(edm_basis *b);
void free_narrow_tables(edm_basis *b);
// the bucket index of the binary search (a no-op for small bases or when it exists); every
// basis gets one when it is created (edm_matvec.hip)
void build_search_index(edm_basis *b);
int run_eigsolve(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int nev, int krylovdim,
                 double tol, int maxiter, uint64_t seed, const double *vec0, double *vals,
                 double *vecs, edm_solve_info *info);
// the same Lanczos iteration on any stream-ordered y = A x (edm_lanczos.cpp)
int run_eigsolve_mv(edm_ctx *ctx, int64_t D, const std::function<void(const double2 *, double2 *)> &mv,
                    int nev, int krylovdim, double tol, int maxiter, uint64_t seed,
                    const double *vec0, double *vals, double *vecs, edm_solve_info *info);
// block Lanczos (edm_lanczos.cpp): mv(nv, x, y) queues y_v = A x_v for v < nv
using BlockMatvec = std::function<void(int, const double2 *const *, double2 *const *)>;
// Krylov-Schur for the nev eigenvalues of smallest real part of a general map; vals holds
// 2 * nev doubles (complex, ascending real part) (edm_lanczos.cpp)
int run_eigsolve_general_mv(edm_ctx *ctx, int64_t D,
                            const std::function<void(const double2 *, double2 *)> &mv, int nev,
                            int krylovdim, double tol, int maxiter, uint64_t seed,
                            const double *vec0, double *vals, double *vecs, edm_solve_info *info);
// complex Schur form A = Z T Z^H, diagonal ascending by real part (interleaved complex,
// row-major n x n; Z column c = Schur vector c)
int projected_schur(int n, const double *A, double *T, double *Z);
int run_eigsolve_block(edm_ctx *ctx, int64_t D, int bs, const BlockMatvec &mv, int nev,
                       int krylovdim, double tol, int maxiter, uint64_t seed, const double *vec0,
                       double *vals, double *vecs, edm_solve_info *info);
int run_eigsolve_block_op(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int bs, int nev,
                          int krylovdim, double tol, int maxiter, uint64_t seed,
                          const double *vec0, double *vals, double *vecs, edm_solve_info *info);
// stored matrix (edm_csr.hip): row keys (column << 32 | term << 1 | sign) of every
// contribution, written at rowptr[r] (edm_matvec.hip)
void launch_csr_keys(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const int64_t *d_rowptr,
                     uint64_t *d_keys);
// time the default matvec kernel at 256/512/1024-row workgroups, and the same with the
// amplitudes staged in LDS when the image has that tail, on this (op, basis); keep the
// fastest for later launches. ms6 gets the six timings (-1: not run); returns the chosen
// rows per workgroup (| EDM_TUNE_AMP_LDS for the LDS-amplitude kernel) or 0
// ms[EDM_TUNE_CANDIDATES]: milliseconds per launch of each candidate (-1: not run)
int matvec_tune(edm_ctx *ctx, const edm_op *op, const edm_basis *b, double *ms);
// eigenvector consumers (edm_matvec.hip)
void launch_matvec_between(edm_ctx *ctx, const edm_op *op, const edm_basis *bra,
                           const edm_basis *ket, const double2 *x, double2 *y);
void launch_rdm1(edm_ctx *ctx, const edm_basis *b, const double2 *psi, const int2 *d_pairs,
                 int npairs, bool lin, double2 *d_out);
// d_out[0] = <bra| op2 op1 |ket> (generic term lists, binary search in the bra basis)
void launch_bracket2(edm_ctx *ctx, const edm_op *op2, const edm_op *op1, const edm_basis *bra,
                     const edm_basis *ket, const double2 *d_bra, const double2 *d_ket,
                     double2 *d_out);
// the device address of the file-scope scratch array on the current device
double2 *noalloc_scratch_address();
// the same with the chunk partials in the library's file-scope scratch array (no device
// allocation, no host synchronisation; d_out is the caller's device complex)
void launch_bracket2_noalloc(edm_ctx *ctx, const edm_op *op2, const edm_op *op1,
                             const edm_basis *bra, const edm_basis *ket, const double2 *d_bra,
                             const double2 *d_ket, double2 *d_out);
// M (nA x B->D, row-major, zeroed by the caller) of the particle reduced density matrix
// rdm = M M^H; d_counts[0] += (state, part) pairs with the part inside the transformed state,
// d_counts[1] += those whose complement is in B
void launch_pes_fill(edm_ctx *ctx, const edm_basis *psi_b, const double2 *d_psi, int64_t nA,
                     const uint64_t *d_A, uint64_t ph_mask, const edm_basis *B, double2 *d_M,
                     unsigned long long *d_counts);
void launch_step_gram(edm_ctx *ctx, const edm_basis *b, const double2 *d_u,
                      const double2 *const *d_vf, const double2 *const *d_vi, int g, double2 *d_out);
struct edm_csr {
    edm_ctx *ctx = nullptr;
    int64_t D = 0, nnz = 0;
    int64_t *d_rowptr = nullptr;   // D + 1
    int32_t *d_cols = nullptr;     // nnz, ascending within a row
    double2 *d_vals = nullptr;     // nnz
};
edm_csr *csr_build(edm_ctx *ctx, const edm_op *op, const edm_basis *b);
void csr_destroy(edm_csr *m);
void csr_matvec(edm_ctx *ctx, const edm_csr *m, const double2 *x, double2 *y);
// A (D x D complex, row-major, device) = the stored matrix with zeros elsewhere
void csr_dense(edm_ctx *ctx, const edm_csr *m, double2 *A);
const char *edm_global_error();
// eigensolver for a small real symmetric matrix (edm_lanczos.cpp)
void sym_eigh(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &Z);
// the same for a complex Hermitian matrix (row-major A; Z column-major)
void herm_eigh(int n, std::vector<std::complex<double>> &A, std::vector<double> &w,
               std::vector<std::complex<double>> &Z);
// batched matvec over many problems (edm_matvec.hip). batch_mv_desc writes the problem's
// descriptor (batch_mv_desc_size() bytes) for y = H x and returns false when the operator has
// no LDS-staged table (the caller then runs launch_matvec for it); blocks are 512 rows.
size_t batch_mv_desc_size();
// workgroups of `block` rows (512 in the shared launches, 1,024 for big groups) over D rows
int64_t batch_mv_blocks(int64_t D, int block);
bool batch_mv_desc(const edm_op *op, const edm_basis *b, const double2 *x, double2 *y,
                   int64_t blk0, void *out, size_t *lds_bytes);
void launch_batch_mv(edm_ctx *ctx, const void *d_descs, const int *d_blk2prob, int64_t nblocks,
                     size_t lds_bytes);
// the shared-walk form for G (2..batch_mv_max_group()) problems on one basis whose operators
// have the same term structure (batch_mv_same_shape): one descriptor covers all G, x[g] and
// y[g] per problem; false when the group cannot share (then run them one by one)
int batch_mv_max_group();
size_t batch_mv_group_desc_size();
bool batch_mv_same_shape(const edm_op *a, const edm_op *b);
// xi: the group's x vectors interleaved by launch_batch_interleave (nullptr: read x[g])
// big: the group has its own launch of 1,024-row workgroups (groups of 2^18 rows and more),
// whose LDS may hold the 14-bit rank split of either form
bool batch_mv_group_desc(const edm_op *const *ops, const edm_basis *b, const double2 *const *x,
                         double2 *const *y, int G, int64_t blk0, void *out, size_t *lds_bytes,
                         const double2 *xi, bool big);
// xi[j * G + g] = x[g][j] for the D rows of G vectors (G = 2..batch_mv_max_group())
void launch_batch_interleave(edm_ctx *ctx, int G, const double2 *const *x, int64_t D,
                             double2 *xi);
void launch_batch_mv_group(edm_ctx *ctx, int G, const void *d_descs, const int *d_blk2grp,
                           int64_t nblocks, size_t lds_bytes, int block);
// multi-problem Lanczos engine (edm_batch.cpp)
edm_batch *batch_create(edm_ctx *ctx);
void batch_destroy(edm_batch *bt);
edm_ctx *batch_ctx(edm_batch *bt);
int batch_add(edm_batch *bt, const edm_op *op, const edm_basis *b, int nev, int krylovdim,
              double tol, int maxiter, uint64_t seed, const double *vec0, int want_vectors,
              int64_t *id);
int batch_step(edm_batch *bt, int max_steps, int *n_running, int *n_finished);
int batch_finished(edm_batch *bt, int64_t *ids, int cap, int *n);
int batch_status(edm_batch *bt, int64_t id, int *state, edm_solve_info *info);
int batch_result(edm_batch *bt, int64_t id, double *vals, double *vecs, edm_solve_info *info);
int batch_remove(edm_batch *bt, int64_t id);
int batch_count(edm_batch *bt, int *n_total, int *n_running);
uint64_t batch_problem_bytes(int64_t D, int nev, int krylovdim);
int batch_set_quantum(edm_batch *bt, int64_t rows);
int batch_set_weight(edm_batch *bt, int64_t id, double weight);
// vector kernels (edm_blas.hip)
void blas_fill_random(edm_ctx *ctx, double2 *v, int64_t D, uint64_t seed, int mode);
// h[i] = <V_i, w> for i < nvec (V given by a device pointer table); result on device
void blas_multidot(edm_ctx *ctx, const double2 *const *d_vptrs, int nvec, const double2 *w,
                   int64_t D, double2 *d_h);
// w -= sum_i V_i h_i (h on device)
void blas_multiaxpy(edm_ctx *ctx, const double2 *const *d_vptrs, int nvec, const double2 *d_h,
                    double2 *w, int64_t D);
// both passes of classical Gram-Schmidt with one re-orthogonalisation: h = V^H w,
// w -= V h, h2 = V^H w, w -= V h2
void blas_cgs2(edm_ctx *ctx, const double2 *const *d_vptrs, int nvec, double2 *w, int64_t D,
               double2 *d_h, double2 *d_h2);
// d_out[0] = ||w||_2
void blas_norm(edm_ctx *ctx, const double2 *w, int64_t D, double *d_out);
// v = w / (*d_beta)
void blas_scale_inv(edm_ctx *ctx, double2 *v, const double2 *w, const double *d_beta, int64_t D);
// batched CGS2 + norm for every problem in d_descs (nprob problems, at most maxvec vectors):
// h = V^H w, w -= V h, h2 = V^H w, w -= V h2, beta = ||w||
void bblas_cgs2_norm(edm_ctx *ctx, const BatchVecDesc *d_descs, int nprob, int maxvec,
                     const int *d_tile2prob, int64_t ntiles, const int *d_blk2prob, int64_t nblk);
// w /= beta for every problem
void bblas_scale(edm_ctx *ctx, const BatchVecDesc *d_descs, const int *d_blk2prob, int64_t nblk);
// batched rotation jobs (blk2job maps each 256-row block to its job)
void bblas_rotate(edm_ctx *ctx, const BatchRotDesc *d_jobs, const int *d_blk2job, int64_t nblk,
                  int maxm);
// out_c = sum_i V_i Y[i + ldy*c] for c < k (Y real, on device)
void blas_rotate(edm_ctx *ctx, const double2 *const *d_vptrs, int m, const double *d_Y, int ldy,
                 int k, double2 *const *d_outptrs, int64_t D);
// out_c = sum_i V_i Y[i + ldy*c] with complex Y (column-major), c < k
void blas_rotate_c(edm_ctx *ctx, const double2 *const *d_vptrs, int m, const double2 *d_Y, int ldy,
                   int k, double2 *const *d_outptrs, int64_t D);
