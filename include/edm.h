/*
 * edm.h — C-ABI of libedm.so, the MI355X (gfx950) exact-diagonalisation engine for
 * momentum-conserved fermion systems.
 *
 * The entry points replace the device plugin contract of MomentumED.jl
 * (MomentumED/src/method/gpu_interface.jl:1-22 and its CUDA implementation
 * MomentumED/ext/MomentumEDCUDAExt.jl). Each declaration names the reference interface
 * it stands in for. Conventions:
 *   - every function returns an int status: EDM_OK (0) or a negative EDM_E* code; the
 *     message is available from edm_last_error(ctx) (or edm_last_error(NULL) for errors
 *     raised before a context exists);
 *   - host buffers are owned by the caller and copied; device handles (edm_basis,
 *     edm_op) are owned by the library until destroyed;
 *   - complex numbers are interleaved (re, im) doubles (Julia ComplexF64 / numpy
 *     complex128 layout);
 *   - basis states are uint64 occupation bitstrings, bit i-1 = orbital i
 *     (EDCore/src/types/manybodystate_basis.jl:24-32,161-168);
 *   - one HIP stream per context; calls on one context are not thread-safe, separate
 *     contexts are independent.
 */
#ifndef EDM_H
#define EDM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDM_OK 0
#define EDM_EINVAL -1    /* bad argument (ArgumentError / AssertionError analogue) */
#define EDM_ENODEV -2    /* no usable GPU (gpu_interface.jl:63-80 _throw_gpu_unavailable) */
#define EDM_ENOMEM -3    /* device allocation failed or HBM budget exceeded */
#define EDM_EHIP -4      /* HIP runtime error */
#define EDM_ENOCONV -5   /* "Krylov method fails. Cannot find N eigenvectors." (MomentumED.jl:466-468) */
#define EDM_EDIM -6      /* DimensionMismatch (linear_map.jl:93-98) */

typedef struct edm_ctx edm_ctx;
typedef struct edm_basis edm_basis;
typedef struct edm_op edm_op;
typedef struct edm_batch edm_batch;
typedef struct edm_csr edm_csr;

/* Description of a momentum-conserved system and particle numbers
 * (EDPara, MomentumED/src/preparation/init_parameter.jl:62-146, and the N_each_component
 * argument of ED_momentum_subspaces, momentum_decomposition.jl:188-192). */
typedef struct {
    int32_t dim;             /* spatial dimension of momenta, 1..3 */
    int32_t nk;              /* Nk = number of momenta (columns of k_list) */
    int32_t nc_mix;          /* Nc_mix (components mixed by H) */
    int32_t nc_conserve;     /* Nc_conserve (components with conserved particle number) */
    const int64_t *k_list;   /* dim x Nk, column-major: k_list[d + dim*k] */
    const int64_t *gk;       /* Gk[dim]; 0 = no periodicity in that direction */
    const int64_t *n_each;   /* particle number per conserved component (>= 0) */
} edm_system;

typedef struct {
    int32_t iterations;      /* restart cycles used */
    int32_t matvecs;         /* operator applications */
    int32_t converged;       /* number of wanted eigenpairs with residual <= tol */
    int32_t krylovdim;       /* Krylov dimension actually used */
    double max_residual;     /* largest residual norm among the returned pairs */
    double seconds;          /* wall time of the solve */
} edm_solve_info;

/* ----------------------------------------------------------------- device / context */
/* GPU_AVAILABLE[:rocm] (gpu_interface.jl:25-30; MomentumEDCUDAExt.jl:13-24):
 * 1 if a gfx950-class device is usable, 0 otherwise. */
int edm_device_available(void);

/* activate_gpu_device + the extension's per-device state (gpu_interface.jl:41-60).
 * hbm_budget_bytes = 0 means no budget beyond the device's memory. */
int edm_ctx_create(int device, uint64_t hbm_budget_bytes, edm_ctx **out);
/* release_gpu_memory(::Val{dev}, level) (gpu_interface.jl:87; MomentumEDCUDAExt.jl:43-47)
 * for everything the context owns. */
int edm_ctx_destroy(edm_ctx *ctx);
const char *edm_last_error(const edm_ctx *ctx);
/* Block until all work queued on the context's stream is complete. */
int edm_synchronize(edm_ctx *ctx);
/* The HIP stream (hipStream_t) the context launches on, for callers sharing device memory. */
int edm_ctx_stream(edm_ctx *ctx, void **stream);
/* Device bytes currently held by the context's handles. */
int edm_ctx_memory(edm_ctx *ctx, uint64_t *used_bytes, uint64_t *budget_bytes);

/* ----------------------------------------------------------------- basis (a1) */
/* ED_momentum_subspaces (momentum_decomposition.jl:188-262): number of momentum sectors in
 * momentum_list order (first coordinate fastest, :209-219), including empty ones. */
int edm_num_sectors(const edm_system *sys, int64_t *nsec);
/* Sector sizes of every sector, enumerated on the device (counts[nsec]). */
int edm_sector_sizes(edm_ctx *ctx, const edm_system *sys, int64_t *counts);
/* One sector's basis (ascending uint64 states) enumerated on the device by colex
 * unranking + momentum filter + order-preserving compaction; `sector` is the index into
 * momentum_list. The handle carries rank tables for partner lookup. */
int edm_basis_enumerate(edm_ctx *ctx, const edm_system *sys, int64_t sector, edm_basis **out);
/* Every sector's basis at once, for ED_momentum_subspaces without momentum_restriction
 * (momentum_decomposition.jl:188-262, which walks all states once on the host): one device
 * counting pass per 64 sectors and one writing pass per 16 sectors over the occupation rank
 * space, instead of a counting and a writing pass per sector. out[k] (k < the number of
 * sectors in momentum_list, edm_sector_sizes order) receives sector k's basis, as
 * edm_basis_enumerate(ctx, sys, k) would return it (also for empty sectors, D = 0). */
int edm_basis_enumerate_all(edm_ctx *ctx, const edm_system *sys, edm_basis **out);
/* HilbertSubspace(list) (EDCore/src/types/hilbert_subspace.jl:19-30) uploaded as-is
 * (must be sorted ascending, unique). `sys` may be NULL; when given and the list is the
 * complete sector `sector`, that sector's rank tables replace the binary search; otherwise
 * rank tables are derived from the list when it splits into equal blocks (rank mode 2,
 * edm_basis_rank_mode). hilbert_subspace.jl:67-72 semantics are kept either way. */
int edm_basis_upload(edm_ctx *ctx, const uint64_t *states, int64_t D, int32_t nbits,
                     const edm_system *sys, int64_t sector, edm_basis **out);
int edm_basis_size(const edm_basis *b, int64_t *D);
int edm_basis_download(const edm_basis *b, uint64_t *states_host);
/* Device pointer to the D sorted states (valid until edm_basis_destroy). */
int edm_basis_device_ptr(const edm_basis *b, const uint64_t **states_dev);
/* How partner indices are found (searchsortedfirst semantics, hilbert_subspace.jl:67-72, in
 * every mode): 0 = binary search over the sorted states (gpu_interface.jl:98-113);
 * 1 = rank tables of the complete sector `sector` of the system given at upload or
 * enumeration; 2 = rank tables derived from the state list alone (edm_basis_upload without a
 * system, or with one whose sector the list is not), which an operator uses only after a
 * device check that every target of every state under it lies in the basis, and binary
 * search otherwise. The reference's device extension uploads bare state lists
 * (MomentumEDCUDAExt.jl:94-113), so mode 2 is what its call sequence gets. Tables (2 x
 * 2^(nbits/2) 32-bit words) are built for up to 52 bits when they take at most 64 MB or four
 * times the basis' 8*D bytes; other bases search (mode 0). */
int edm_basis_rank_mode(const edm_basis *b, int32_t *mode);
/* Release the rank tables (mode 0 from then on). No reference counterpart: it lets a caller
 * compare the two lookups or give back the tables' memory. Not while a batch holds a
 * problem on this basis (its step plans point at the tables). */
int edm_basis_drop_rank(edm_basis *b);
int edm_basis_destroy(edm_basis *b);

/* ----------------------------------------------------------------- operator (a2-a5) */
/* create_gpu_linearmap(A::LinearMap, ::Val{dev}) (gpu_interface.jl:85;
 * MomentumEDCUDAExt.jl:94-113) from an MBOperator's term list
 * (EDCore/src/types/operator.jl:24-32): S terms with amplitude amp[2S] (interleaved),
 * creation mask out[S] and annihilation mask in[S]. When upper_hermitian != 0 the list is
 * the upper triangle and the adjoint of every non-diagonal term is added
 * (MomentumED/src/method/linear_map.jl:22-41). */
int edm_op_upload(edm_ctx *ctx, int32_t nbits, const double *amp, const uint64_t *out,
                  const uint64_t *in, int64_t S, int32_t upper_hermitian, edm_op **op);
/* Number of terms in the expanded (full) list. */
int edm_op_size(const edm_op *op, int64_t *S_full);
/* The expanded list in (in, out) order, as linear_map.jl:22-41 sorts it. Arrays of
 * length S_full; parity = parity_mask(out) xor parity_mask(in) (scattering.jl:7-18,60). */
int edm_op_full_list(const edm_op *op, double *amp, uint64_t *out, uint64_t *in,
                     uint64_t *parity);
int edm_op_destroy(edm_op *op);

/* ----------------------------------------------------------------- matvec (a6-a8) */
/* (A::CuLinearMap)(y, x) (MomentumEDCUDAExt.jl:119-122 -> gpu_interface.jl:214-220),
 * computing y = H x with the semantics of the CPU map (linear_map.jl:100-118); targets
 * outside the basis are dropped. x_dev, y_dev are device pointers to D complex values
 * (16-byte aligned); stream-ordered on the context's stream (no host sync). */
int edm_matvec(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const void *x_dev,
               void *y_dev);
/* The adjoint map (AdjointLinearMap, linear_map.jl:120-139): y = H^dagger x. */
int edm_matvec_adjoint(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const void *x_dev,
                       void *y_dev);
/* Block form of the map: y_v = H x_v for v < nv (the map applied to nv vectors, as a block
 * Krylov method needs it; the reference applies its map one vector at a time,
 * gpu_interface.jl:214-220). x_dev[v], y_dev[v] are device pointers to D complex values;
 * work_dev is caller-owned device scratch of min(nv, 8) * D complex values (it receives an
 * interleaved copy of the x_v so that one partner's nv entries are contiguous). Column v
 * equals edm_matvec(x_dev[v]) bit for bit. Stream-ordered. */
int edm_matvec_block(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int nv,
                     const void *const *x_dev, void *const *y_dev, void *work_dev);
/* Host-buffer convenience: copies x in, runs edm_matvec, copies y out, synchronises. */
int edm_matvec_host(edm_ctx *ctx, const edm_op *op, const edm_basis *b, const double *x,
                    double *y);

/* Connectivity of rows [row0, row0+nrows): which (row, column, term, sign) pairs contribute
 * (linear_map.jl:109-113). counts[nrows] first; then, with rowptr[nrows+1] = exclusive prefix
 * of counts, cols[], terms[] (index into edm_op_full_list order) and negs[] (1 = sign -1).
 * Entries of a row are sorted by (column, term). Host buffers. */
int edm_connectivity_counts(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t row0,
                            int64_t nrows, int64_t *counts);
int edm_connectivity(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t row0,
                     int64_t nrows, const int64_t *rowptr, int64_t *cols, int32_t *terms,
                     int8_t *negs);
/* Totals of the same connectivity over the whole basis, counted on the device: all
 * contributing (row, term) pairs and the off-diagonal ones (target != row state). The
 * off-diagonal count is the number of x gathers one matvec performs (the nnz_off of the
 * algorithmic-bytes model). No reference counterpart: it sizes the roofline of a6. */
int edm_connectivity_totals(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int64_t *nnz,
                            int64_t *nnz_offdiag);

/* Autotuner (no reference counterpart; the reference has one GPU kernel configuration): times
 * on this operator and basis (three launches each, on scratch vectors) the default matvec
 * kernel with 256-, 512- and 1024-row workgroups, the variant that also stages the (distinct)
 * amplitudes in LDS at the same three sizes when they fit, the split-row variant with 2, 4, 8
 * and 16 lanes per row (512-lane workgroups), and, when the basis has a rank split whose low
 * table (at most 2^14 16-bit entries) fits in LDS beside the term image (built at upload
 * beside the rank tables of complete sectors: a dense high table up to 38 bits when it stays
 * within the basis' own size, a two-level one from 39 to 50 bits for 2^18 states or more),
 * the variant that reads the low-half rank table from LDS (and gathers x
 * into LDS while its walk goes on) with 512- and 1024-row workgroups; it uses the
 * fastest for every later edm_matvec / edm_eigsolve on the pair. The per-row arithmetic and
 * its order are the same in all twelve, so y does not change. *kernel = chosen code: rows per
 * workgroup, | EDM_TUNE_AMP_LDS for the LDS-amplitude variant, | EDM_TUNE_NARROW for the
 * LDS rank-table variant, | lanes << EDM_TUNE_SPLIT_SHIFT for the split-row variant (0 when
 * the pair runs another kernel); ms (may be NULL) = ms[EDM_TUNE_CANDIDATES], the mean
 * milliseconds per launch of 256, 512, 1024, LDS-amplitude 256, 512, 1024, split 2, 4, 8, 16,
 * LDS-rank 512, 1024 (-1: not run). A pair that was not tuned uses the LDS-rank variant with
 * 1024-row workgroups when D >= 2^18 and the basis has such a split, else one lane per row
 * when D >= 2^18, and otherwise the smallest power of two of lanes (at most 4) that brings
 * D * lanes to 2^18. */
#define EDM_TUNE_AMP_LDS 0x10000
#define EDM_TUNE_NARROW 0x20000
#define EDM_TUNE_SPLIT_SHIFT 20
#define EDM_TUNE_CANDIDATES 12
int edm_matvec_tune(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t *kernel,
                    double *ms);
/* The kernel code (as edm_matvec_tune reports it) that edm_matvec(op, b) launches now: the
 * tuner's pick or the size-based default, 0 when the pair runs the global-table kernels. No
 * reference counterpart: it names the measured kernel. */
int edm_matvec_kernel(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t *kernel);

/* ----------------------------------------------------------------- eigensolve (a9-a10) */
/* krylov_map_solve(H::AbstractGPULinearMap, N; ishermitian=true, vec0, krylovdim, tol,
 * maxiter) (gpu_interface.jl:92-94; MomentumEDCUDAExt.jl:159-179 -> KrylovKit.eigsolve(...,
 * :SR)): the nev lowest eigenpairs by thick-restart Lanczos in complex FP64 with full
 * re-orthogonalisation on the device. vec0 (host, D complex) may be NULL: then a seeded
 * uniform [0,1)+i[0,1) start vector is used. vals[nev] ascending; vecs (host, nev*D
 * complex, vector-major) may be NULL. Returns EDM_ENOCONV when fewer than nev pairs reach
 * tol within maxiter restarts (vals/vecs still hold the best estimates). */
int edm_eigsolve(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t nev,
                 int32_t krylovdim, double tol, int32_t maxiter, uint64_t seed,
                 const double *vec0, double *vals, double *vecs, edm_solve_info *info);

/* krylov_map_solve(H, N; ishermitian=false) (linear_map.jl:179-191; MomentumEDCUDAExt.jl:
 * 159-179 -> KrylovKit.eigsolve(..., :SR; ishermitian=false)): the nev eigenvalues of smallest
 * real part of a general (non-Hermitian) operator's map, by Krylov-Schur (Arnoldi with full
 * re-orthogonalisation and restarts that keep the leading Schur vectors of the projected
 * matrix; Stewart 2001). vals[2*nev] complex (re, im), ascending real part; vecs (host,
 * nev*D complex, unit-norm right eigenvectors) may be NULL. Convergence, arguments and
 * returns as edm_eigsolve (absolute residual |b . y_i| <= tol). */
int edm_eigsolve_general(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t nev,
                         int32_t krylovdim, double tol, int32_t maxiter, uint64_t seed,
                         const double *vec0, double *vals, double *vecs, edm_solve_info *info);

/* The same eigenpairs by thick-restart block Lanczos with block size `block` (1..8): each
 * step applies the map to `block` vectors at once (edm_matvec_block) and orthonormalises
 * them against the whole basis (CGS2, column by column); the projected matrix is complex
 * Hermitian. krylovdim is the basis size (rounded up to a multiple of block). Fewer host
 * round trips per matvec, one shared candidate walk per block, and robust on degenerate
 * multiplets up to `block` levels; there is no counterpart in the reference (KrylovKit's
 * Lanczos is single-vector). block = 1, or a subspace with D < krylovdim + block, runs
 * edm_eigsolve's iteration. Arguments and returns as edm_eigsolve (vec0 seeds the first
 * block column; the others are seeded random vectors). */
int edm_eigsolve_block(edm_ctx *ctx, const edm_op *op, const edm_basis *b, int32_t nev,
                       int32_t block, int32_t krylovdim, double tol, int32_t maxiter,
                       uint64_t seed, const double *vec0, double *vals, double *vecs,
                       edm_solve_info *info);

/* ----------------------------------------------------------------- eigenvector consumers */
/* y = O x between two bases: y_i = sum_j <bra_i|O|ket_j> x_j (y has bra's length, x ket's),
 * the product inside ED_bracket(bra, op, ket) (EDCore/src/multiplication.jl:153-307) when the
 * two vectors live in different subspaces (e.g. an operator that moves momentum). With
 * bra == ket it is edm_matvec. One- and two-body number-conserving terms. */
int edm_matvec_between(edm_ctx *ctx, const edm_op *op, const edm_basis *bra,
                       const edm_basis *ket, const void *x_dev, void *y_dev);
/* ED_bracket(bra, op2, op1, ket) / ED_bracket_threaded (EDCore/src/multiplication.jl:
 * 170-183, 214-223, 286-307): <bra| op2 op1 |ket> with both operators' full term lists
 * (an upper-Hermitian operator includes the adjoints of its off-diagonal terms, as
 * hermitian_complete does), applied from the left to every ket state in turn; final states
 * outside the bra basis are dropped. bra and ket may be different subspaces (e.g. a density
 * operator pair rho(-q) rho(q)); no intermediate subspace is needed. Any term shape.
 * out[0..1] = the complex result. */
int edm_bracket2(edm_ctx *ctx, const edm_op *op2, const edm_op *op1, const edm_basis *bra,
                 const edm_basis *ket, const void *bra_dev, const void *ket_dev, double *out);
/* edm_bracket2 for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * The same kernels and summation order as edm_bracket2, so the same result bit for bit, but
 * stream-ordered: out_dev (the caller's device memory, one complex) receives
 * <bra| op2 op1 |ket> when the stream reaches it, with no host synchronisation. The
 * partial sums live in a file-scope device array of the library (4096 complex values, 64 KB
 * per device, held until the process ends); the bracket needs at most 2048 of them, and a
 * call that would need more is rejected with EDM_EINVAL. It calls no hipMalloc, hipFree or
 * allocating library routine. Every call on one device shares that array: calls that overlap
 * on one device (two streams, or two host threads) race. */
int edm_bracket2_noalloc(edm_ctx *ctx, const edm_op *op2, const edm_op *op1,
                         const edm_basis *bra, const edm_basis *ket, const void *bra_dev,
                         const void *ket_dev, void *out_dev);
/* PES_1rdm (MomentumED/src/analysis/particle_reduced_density_matrix.jl:4-33): for each of
 * npairs (i, j) (0-based orbitals, pairs[2p], pairs[2p+1]) out[2p..2p+1] = <psi|c_i^+ c_j|psi>
 * for the device vector psi in basis b. On a complete sector, pairs that do not conserve it
 * are exactly 0 and are not computed. */
int edm_rdm1(edm_ctx *ctx, const edm_basis *b, const void *psi_dev, int32_t npairs,
             const int32_t *pairs, double *out);
/* PES_MomtBlock_rdm (MomentumED/src/analysis/particle_reduced_density_matrix.jl:37-88)
 * factorised: the particle reduced density matrix of psi on the part-A block (A: nA host
 * states) is rdm = sum over complement blocks B of M M^H, with M[a, b] = (-1)^occ(s, A[a])
 * psi[s] for s ^ ph_mask = A[a] | B[b] (occ = scat_occ_number of the untransformed state,
 * EDCore/src/types/manybodystate_basis.jl:397-409; ph_mask = the bits of the particle-hole
 * transformed components). This writes M (nA x D_B complex, row-major, device, zeroed by
 * the caller) for psi (device vector in psi_basis) and one complement block B; the product
 * is one dense GEMM. counts NULL: every (state, part) pair with the part inside the
 * transformed state must find its complement in B (EDM_EINVAL otherwise). counts non-NULL:
 * counts[0] = such pairs, counts[1] = those found in B, the others skipped, so that the
 * caller can split the complements over several blocks and check that their counts[1] add
 * up to counts[0]. */
int edm_pes_block_matrix(edm_ctx *ctx, const edm_basis *psi_basis, const void *psi_dev,
                         int64_t nA, const uint64_t *A, uint64_t ph_mask, const edm_basis *B,
                         void *m_dev, int64_t *counts);
/* ED_step_inner_prod (MomentumED/src/analysis/manybody_quantum_geometry.jl:43-91): for g
 * vectors each (device pointers vf_dev[a], vi_dev[b] in basis b) and orbital inner products
 * u[nbits] (host, complex), out[a*g+b] = sum_x conj(f_a[x]) i_b[x] prod_{k occupied in x} u[k]. */
int edm_step_inner_prod(edm_ctx *ctx, const edm_basis *b, int32_t g, const void *const *vf_dev,
                        const void *const *vi_dev, const double *orbital_ip, double *out);

/* ----------------------------------------------------------------- stored matrix (:sparse) */
/* SparseHmltMatrix + create_gpu_matrix (MomentumED/src/method/sparse_matrix.jl:65-191;
 * MomentumEDCUDAExt.jl:147-152): the operator's matrix in the basis, built on the device and
 * kept there as CSR (int64 row offsets, int32 column indices ascending within each row,
 * complex128 values; entries of equal (row, column) summed, as the reference's CSC build
 * does). Row i holds <i|H|j>. One- and two-body number-conserving operators; D < 2^31 and at
 * most 2^31 contributions. */
int edm_csr_build(edm_ctx *ctx, const edm_op *op, const edm_basis *b, edm_csr **out);
int edm_csr_size(const edm_csr *m, int64_t *D, int64_t *nnz);
/* Host copies: rowptr[D+1], cols[nnz], vals[2*nnz]. */
int edm_csr_download(const edm_csr *m, int64_t *rowptr, int32_t *cols, double *vals);
/* y = A x on device vectors (the CuSparseMatrixCSR product inside krylov_matrix_solve),
 * stream-ordered. */
int edm_csr_matvec(edm_ctx *ctx, const edm_csr *m, const void *x_dev, void *y_dev);
/* krylov_matrix_solve(H_gpu, N; ishermitian=true, ...) (method/gpu_interface.jl:92-94 for
 * matrices; MomentumED.jl:316-333): arguments and results as edm_eigsolve. */
int edm_csr_eigsolve(edm_ctx *ctx, const edm_csr *m, int32_t nev, int32_t krylovdim, double tol,
                     int32_t maxiter, uint64_t seed, const double *vec0, double *vals,
                     double *vecs, edm_solve_info *info);
/* krylov_matrix_solve(H_gpu, N; ishermitian=false) (sparse_matrix.jl:248-255): the general
 * eigensolver of edm_eigsolve_general on the stored matrix. */
int edm_csr_eigsolve_general(edm_ctx *ctx, const edm_csr *m, int32_t nev, int32_t krylovdim,
                             double tol, int32_t maxiter, uint64_t seed, const double *vec0,
                             double *vals, double *vecs, edm_solve_info *info);
/* Matrix(H) of the reference's :dense method (MomentumED.jl:340-369, which densifies the
 * sparse matrix and calls LAPACK on the host): the stored matrix written as a dense D x D
 * complex row-major array at a_dev (device, 16 * D * D bytes, entry (i, j) = <i|H|j>), for a
 * dense Hermitian eigensolver on the device. Stream-ordered. */
int edm_csr_dense(edm_ctx *ctx, const edm_csr *m, void *a_dev);
int edm_csr_destroy(edm_csr *m);

/* ----------------------------------------------------------------- batched / served solves */
/* The reference has no batch object: a sweep over sectors and parameter points is a host loop
 * of EDsolve calls (MomentumED.jl:375-480; e.g. the per-sector loop in
 * examples/landau level torus/topological order/4. cuda support.ipynb, cell source lines
 * 299-301). A batch advances many independent krylov_map_solve
 * problems together: each problem is exactly the solve edm_eigsolve runs (same arguments,
 * same iteration, bitwise the same eigenvalues and vectors), but one step moves every running
 * problem by one Lanczos iteration with one launch per kernel for all of them. Problems can be
 * added and collected between steps. The operator and basis of a problem must stay alive
 * until the problem is collected or removed. A batch lives on its context's stream. */
int edm_batch_create(edm_ctx *ctx, edm_batch **out);
int edm_batch_destroy(edm_batch *bt);
/* Admit a problem (arguments as edm_eigsolve; want_vectors keeps the nev eigenvectors on the
 * device for edm_batch_result). Takes the problem's workspace (edm_batch_problem_bytes) from
 * the context: with an HBM budget as one allocation of its own (EDM_ENOMEM when the budget
 * does not hold it), without one as a range of a 1 GiB chunk the batch shares between its
 * problems (a chunk returns to the device once no problem uses it). *id names it. */
int edm_batch_add(edm_batch *bt, const edm_op *op, const edm_basis *b, int32_t nev,
                  int32_t krylovdim, double tol, int32_t maxiter, uint64_t seed,
                  const double *vec0, int32_t want_vectors, int64_t *id);
/* Run up to max_steps Lanczos steps of all running problems, returning early after a step in
 * which any problem finished. *n_running: problems still running; *n_finished: problems that
 * finished during this call (either pointer may be NULL). */
int edm_batch_step(edm_batch *bt, int32_t max_steps, int32_t *n_running, int32_t *n_finished);
/* Ids of finished, not yet collected problems (up to cap written; *n = total count). */
int edm_batch_finished(edm_batch *bt, int64_t *ids, int32_t cap, int32_t *n);
/* *state: 0 running, 1 converged, 2 finished without all nev pairs converged; info may be NULL. */
int edm_batch_status(edm_batch *bt, int64_t id, int32_t *state, edm_solve_info *info);
/* Collect a finished problem: vals[nev], vecs (host, nev*D complex, vector-major; NULL or the
 * problem must have been added with want_vectors), info; releases its workspace. Returns
 * EDM_OK or EDM_ENOCONV as edm_eigsolve would have. */
int edm_batch_result(edm_batch *bt, int64_t id, double *vals, double *vecs, edm_solve_info *info);
/* Drop a problem, running or finished, and release its workspace. */
int edm_batch_remove(edm_batch *bt, int64_t id);
/* Problems held (running + finished, not collected) and running. */
int edm_batch_count(edm_batch *bt, int32_t *n_total, int32_t *n_running);
/* Step scheduling. rows = 0 (default): every running problem takes part in every step.
 * rows > 0: deficit round robin over rows; each step credits every running problem with rows,
 * and a problem of dimension D takes part when its credit reaches D (then pays D), so small
 * problems step often while large ones step about once per D / rows steps. Only the pace
 * changes; each problem's iteration and result do not. */
int edm_batch_set_quantum(edm_batch *bt, int64_t rows);
/* Weight of one problem in that round robin (default 1): its credit per step is
 * weight * rows, so it steps weight times as often as a problem of the same dimension and
 * weight 1. A server uses it to reserve device time for a large problem that would
 * otherwise step only once every D / rows steps. Pace only; the iteration is unchanged. */
int edm_batch_set_weight(edm_batch *bt, int64_t id, double weight);
/* Device bytes of the workspace edm_batch_add takes for a problem of dimension D. */
int edm_batch_problem_bytes(int64_t D, int32_t nev, int32_t krylovdim, uint64_t *bytes);

/* Host eigen-decomposition of a real symmetric n x n matrix (row-major A, not modified):
 * w[n] ascending, Z[n*n] with eigenvector i in Z[i*n .. i*n+n). The Rayleigh-Ritz step of
 * every Lanczos restart (edm_eigsolve, batches) uses it on the projected matrix; KrylovKit
 * does the same with LAPACK (the Lanczos factorization's eigen-decomposition inside
 * eigsolve). No device work. */
int edm_projected_eigh(int32_t n, const double *A, double *w, double *Z);
/* The same for a complex Hermitian n x n matrix (the projected matrix of block Lanczos,
 * edm_eigsolve_block): A and Z hold n*n interleaved (re, im) pairs, A row-major, Z with
 * eigenvector i in Z[2*i*n .. 2*i*n + 2*n). */
int edm_projected_heigh(int32_t n, const double *A, double *w, double *Z);

/* Host complex Schur decomposition A = Z T Z^H of a general n x n matrix (A, T and Z hold
 * n*n interleaved (re, im) pairs, row-major; column c of Z is Schur vector c; the diagonal
 * of T ascending by real part, then imaginary part): the Rayleigh-Ritz step of
 * edm_eigsolve_general, where KrylovKit uses LAPACK's Schur routines. No device work. */
int edm_projected_schur(int32_t n, const double *A, double *T, double *Z);

/* ----------------------------------------------------------------- device vector helpers */
/* Fill D complex values at a device pointer: mode 0 = uniform [0,1) in re and im
 * (rand(ComplexF64), linear_map.jl:187), mode 1 = standard complex normal. Deterministic in
 * (seed, index). */
int edm_fill_random(edm_ctx *ctx, void *v_dev, int64_t D, uint64_t seed, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif /* EDM_H */
