"""Block matvec and block Lanczos on the device, through the C-ABI.

Bars: every column of the block matvec equals the single-vector matvec of that column bit for
bit (same per-row arithmetic and summation order); block Lanczos eigenvalues within 1e-10
relative of dense LAPACK / ARPACK and of the reference's printed energies (which include
exact doublets), eigenvector residuals < 1e-9, orthonormal vectors.
"""
import ctypes
import json
import os

import numpy as np
import pytest
import scipy.sparse as sps

import momentumed as me
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import check

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_printed_values.json")))


def _system(para, N_each):
    return dev.system_struct(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, N_each)


def _oracle_full(O, op):
    amp, out, inn = op.arrays()
    return O.full_list(O.TermList(amp, out, inn), op.upper_hermitian)


def _block_vs_single(ctx, dop, basis, nvs, seed):
    import torch
    D = basis.D
    nmax = max(nvs)
    g = torch.Generator().manual_seed(seed)
    X = [torch.complex(torch.randn(D, generator=g, dtype=torch.float64),
                       torch.randn(D, generator=g, dtype=torch.float64)).cuda()
         for _ in range(nmax)]
    single = []
    for v in range(nmax):
        y = torch.empty_like(X[v])
        torch.cuda.synchronize()
        check(ctx.L.edm_matvec(ctx.h, dop.h, basis.h, X[v].data_ptr(), y.data_ptr()), ctx.h)
        ctx.synchronize()
        single.append(y)
    work = torch.empty(min(nmax, 8) * D, dtype=torch.complex128, device=X[0].device)
    for nv in nvs:
        Y = [torch.full_like(X[0], float("nan")) for _ in range(nv)]
        xp = (ctypes.c_void_p * nv)(*[X[v].data_ptr() for v in range(nv)])
        yp = (ctypes.c_void_p * nv)(*[Y[v].data_ptr() for v in range(nv)])
        torch.cuda.synchronize()
        check(ctx.L.edm_matvec_block(ctx.h, dop.h, basis.h, nv, xp, yp, work.data_ptr()), ctx.h)
        ctx.synchronize()
        for v in range(nv):
            assert torch.equal(Y[v], single[v]), (nv, v)


def test_matvec_block_bitwise_medium(gpu_ctx):
    print("[block] medium 3x9 N=9 sector (0,0): nv 1..11 vs single matvecs", flush=True)
    w = W.medium()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, w.operator(), w.para.bits)
    sector = me.preparation._momentum_list(w.para, w.N_each).index(w.sector)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), sector)
    assert b.rank_mode != 0
    _block_vs_single(gpu_ctx, dop, b, [1, 2, 3, 4, 5, 8, 11], 1)


def test_matvec_block_bitwise_search_and_two_components(gpu_ctx, oracle_mod):
    """Binary-search lookups on a partial basis (dropped targets), a two-component sector,
    and the generic 3-body path (which runs one single matvec per column)."""
    w = W.small()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    lists, _ = oracle_mod.enumerate_sectors(w.para.k_list, w.para.Gk, 1, 1, [5])
    part = lists[0][::3].copy()
    b = dev.DeviceBasis.upload(gpu_ctx, part, w.para.bits, _system(w.para, w.N_each), 0)
    assert b.rank_mode == 2   # derived tables; the closure check sends lookups to search
    _block_vs_single(gpu_ctx, dop, b, [1, 4, 7], 2)
    para = W.batched_para(0.7)
    op2 = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sp2 = me.ED_momentum_subspaces(para, (3, 3))[0][4]
    H2 = dev.create_gpu_linearmap(op2, sp2)
    _block_vs_single(gpu_ctx, H2.op, H2.basis, [4, 8], 3)
    rng = np.random.default_rng(13)
    bits = 12
    terms = []
    for _ in range(40):
        o = rng.choice(bits, 3, replace=False) + 1
        i = rng.choice(bits, 3, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), *map(int, o), *map(int, i),
                                bits=bits, upper_hermitian=True))
    op3 = me.MBOperator(terms, upper_hermitian=True)
    dop3 = dev.DeviceOperator.from_mboperator(gpu_ctx, op3, bits)
    basis3 = np.array([s for s in range(1 << bits) if bin(s).count("1") == 5], dtype=np.uint64)
    b3 = dev.DeviceBasis.upload(gpu_ctx, basis3, bits)
    _block_vs_single(gpu_ctx, dop3, b3, [3], 4)


def test_matvec_block_rejects_aliasing(gpu_ctx):
    import torch
    from momentumed._lib import EDMError
    w = W.small()
    H = dev.create_gpu_linearmap(w.operator(), me.ED_momentum_subspaces(w.para, 5)[0][0])
    x = torch.zeros(H.basis.D, dtype=torch.complex128, device="cuda:0")
    work = torch.zeros(2 * H.basis.D, dtype=torch.complex128, device="cuda:0")
    xp = (ctypes.c_void_p * 2)(x.data_ptr(), x.data_ptr())
    with pytest.raises(EDMError):
        check(gpu_ctx.L.edm_matvec_block(gpu_ctx.h, H.op.h, H.basis.h, 2, xp, xp,
                                         work.data_ptr()), gpu_ctx.h)


def _check_vectors(H, vals, vecs):
    V = np.array(vecs)
    for i in range(len(vals)):
        assert np.linalg.norm(H @ V[i] - vals[i] * V[i]) < 1e-9
    assert np.abs(V.conj() @ V.T - np.eye(len(vals))).max() < 1e-10


def test_block_eigsolve_small_all_sectors(gpu_ctx, oracle_mod):
    print("[block] eigsolve block 4, small 3x5 N=5, all 15 sectors", flush=True)
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    subspaces, _ = me.ED_momentum_subspaces(w.para, 5)
    for sp in subspaces:
        E, V = me.EDsolve(sp, op, N=w.nev, method="map", device="gpu", method_info=False,
                          blocksize=4)
        H = oracle_mod.dense_matrix(sp.list, full)
        ev = np.linalg.eigvalsh(H)
        np.testing.assert_allclose(E, ev[: w.nev], rtol=1e-10, atol=1e-12)
        _check_vectors(H, E, [v.vec for v in V])


@pytest.mark.parametrize("bs", [2, 4, 8])
def test_block_eigsolve_reference_doublets(gpu_ctx, bs):
    """The printed Laughlin energies of sector (0,0) hold exact doublets (the triangular
    lattice's point group); block Lanczos resolves both members of each."""
    import llt
    g = GOLDEN["energies_per_Nk"][0]
    Vi = llt.LandauInteraction(llt.ReciprocalLattice.named("triangular"), (1, 0, 1, 0))
    Vi.D_l, Vi.mix = g["D_l"], g["mix"]
    para = me.EDPara(k_list=W.mesh(*g["mesh"]), Gk=tuple(g["Gk"]), H_two=Vi)
    subspaces, ks = me.ED_momentum_subspaces(para, g["N"])
    H = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sp = subspaces[ks.index(tuple(g["sector"]))]
    E, _, info = dev.krylov_map_solve(dev.create_gpu_linearmap(H, sp), 10, blocksize=bs,
                                      want_vectors=False)
    assert info["converged"] >= 10
    np.testing.assert_allclose(E / para.Nk, g["values"], rtol=1e-10, atol=0)


def test_block_eigsolve_medium_against_arpack(gpu_ctx, oracle_mod):
    import scipy.sparse.linalg as spl
    print("[block] eigsolve block 4, medium 3x9 N=9 sector (0,0)", flush=True)
    w = W.medium()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    sp = me.ED_momentum_subspaces(w.para, 9, momentum_restriction=True,
                                  momentum_list=[(0, 0)])[0][0]
    # A block step raises the Krylov polynomial degree by one for `blocksize` matvecs, so
    # block Lanczos needs a larger basis than the single-vector iteration for the same
    # convergence: at krylovdim 40 (ten blocks of 4) the sixth pair is still at 5e-12 after
    # 100 restarts; at 80 all six reach tol.
    H = dev.create_gpu_linearmap(op, sp)
    E, Vh, info = dev.krylov_map_solve(H, w.nev, krylovdim=2 * w.krylovdim, tol=w.tol,
                                       blocksize=4)
    assert info["converged"] == w.nev, info
    assert info["max_residual"] <= w.tol
    V = [me.MBS64Vector(v, sp) for v in Vh]
    basis = sp.list
    rp, cols, terms, negs = oracle_mod.connectivity(basis, full)
    vals = full.amp[terms] * np.where(negs, -1.0, 1.0)
    Hs = sps.csr_matrix((vals, cols, rp), shape=(len(basis), len(basis)))
    ref = np.sort(spl.eigsh(Hs, k=w.nev, which="SA", tol=1e-14, ncv=60)[0])
    np.testing.assert_allclose(E, ref, rtol=1e-10, atol=0)
    Em, _ = me.EDsolve(sp, op, N=w.nev, method="map", device="gpu", method_info=False,
                       krylovdim=w.krylovdim, tol=w.tol)
    np.testing.assert_allclose(E, Em, rtol=1e-10, atol=0)
    for i, v in enumerate(V):
        r = oracle_mod.matvec(basis, full, v.vec) - E[i] * v.vec
        assert np.linalg.norm(r) < 1e-9
    G = np.array([[np.vdot(a.vec, b.vec) for b in V] for a in V])
    assert np.abs(G - np.eye(len(V))).max() < 1e-10


def test_block_eigsolve_seeded_vec0_and_small_space(gpu_ctx):
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][0]
    H = dev.create_gpu_linearmap(op, sp)
    v1, _, i1 = dev.krylov_map_solve(H, 4, seed=7, blocksize=4)
    v2, _, _ = dev.krylov_map_solve(H, 4, seed=7, blocksize=4)
    np.testing.assert_array_equal(v1, v2)          # deterministic for a fixed seed
    v3, _, _ = dev.krylov_map_solve(H, 4, vec0=np.ones(len(sp)) + 0.5j, blocksize=4)
    np.testing.assert_allclose(v3, v1, rtol=1e-11, atol=1e-12)
    assert i1["converged"] >= 4 and i1["max_residual"] <= 1e-12
    assert i1["krylovdim"] % 4 == 0
    # a subspace smaller than krylovdim + block runs the single-vector iteration: same result
    part = me.HilbertSubspace(sp.list[:30], bits=sp.bits)
    Hp = dev.create_gpu_linearmap(op, part)
    a, _, _ = dev.krylov_map_solve(Hp, 3, krylovdim=30, blocksize=4)
    b, _, _ = dev.krylov_map_solve(Hp, 3, krylovdim=30)
    np.testing.assert_array_equal(a, b)


def test_block_eigsolve_argument_errors(gpu_ctx):
    from momentumed._lib import EDMError
    w = W.small()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][0]
    H = dev.create_gpu_linearmap(w.operator(), sp)
    with pytest.raises(EDMError):
        dev.krylov_map_solve(H, 4, blocksize=9)
    with pytest.raises(NotImplementedError):
        me.EDsolve(sp, w.operator(), N=4, method="sparse", blocksize=4)


@pytest.mark.parametrize("how", ["sparse", "map", "map_block4"])
def test_cdw_reference_energies_with_fourfold_level(gpu_ctx, how):
    """1.CDW.ipynb (n = 2 Landau level, 6x4 mesh, Ne = 6): the 10 printed E/Nk of momentum
    (0, 1), which hold a two-fold and a four-fold level, from the device solvers."""
    import llt
    g = GOLDEN["cdw"]
    V = llt.LandauInteraction(llt.ReciprocalLattice(*g["lattice_ratio_cos"]),
                              *[tuple(c) for c in g["components"]])
    V.D_l, V.mix = g["D_l"], g["mix"]
    para = me.EDPara(k_list=np.array(g["k_list"], dtype=np.int64), Gk=tuple(g["Gk"]), H_two=V)
    subspaces, ks = me.ED_momentum_subspaces(para, g["N"])
    sp = subspaces[ks.index(tuple(g["sector"]))]
    H = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    kw = {"method": "sparse"} if how == "sparse" else {"method": "map"}
    if how == "map_block4":
        kw["blocksize"] = 4
    E, V_ = me.EDsolve(sp, H, N=10, method_info=False, **kw)
    np.testing.assert_allclose(E / para.Nk, g["energies_per_Nk"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("level", ["0", "1"])
def test_pfaffian_cfl_reference_energies(gpu_ctx, oracle_mod, level):
    """3.Pfaffin_vs_CFL.ipynb: the half-filling energies of the n = 0 and n = 1 Landau levels
    (3x6 triangular mesh, Ne = 9). n = 0: the printed values from the default device solve.
    n = 1: the exact spectrum holds a four-fold level of which the reference printed three
    copies; block Lanczos (b = 4) returns the lowest 10 exact values (all four copies) and
    every value of the default single-vector solve is an exact eigenvalue."""
    import llt
    g = GOLDEN["pfaffian_cfl"]
    gl = g["levels"][level]
    V = llt.LandauInteraction(llt.ReciprocalLattice.named(g["lattice"]), (1, int(level), 1, 0))
    V.D_l, V.mix = g["D_l"], g["mix"]
    para = me.EDPara(k_list=np.array(g["k_list"], dtype=np.int64), Gk=tuple(g["Gk"]), H_two=V)
    subspaces, ks = me.ED_momentum_subspaces(para, g["N"])
    sp = subspaces[ks.index(tuple(gl["sector"]))]
    H = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    if level == "0":
        E, _ = me.EDsolve(sp, H, N=10, method_info=False)
        np.testing.assert_allclose(E / para.Nk, gl["energies_per_Nk"], rtol=1e-10, atol=0)
        return
    import test_oracle as T
    exact = T.pfaffian_exact(oracle_mod, level)
    # The default single-vector iteration (the stored-matrix method) converges ten pairs but
    # sees only three copies of the four-fold level: every value it returns is an exact
    # eigenvalue, the list is a submultiset of the exact spectrum.
    Hm = dev.create_gpu_matrix(H, sp)
    E, _, info = dev.krylov_matrix_solve(Hm, 10, want_vectors=False)
    Hm.close()
    assert info["converged"] == 10, info
    T._match_submultiset(E / para.Nk, exact)
    # block Lanczos (b = 4) finds all four copies of the level. Its basis must be larger
    # than the single-vector default: at krylovdim 30 (seven blocks and a half) only 7 of
    # the 10 pairs reach tol within 100 restarts.
    Eb, _, ib = dev.krylov_map_solve(dev.create_gpu_linearmap(H, sp), 10, blocksize=4,
                                     krylovdim=60, want_vectors=False)
    assert ib["converged"] == 10, ib
    np.testing.assert_allclose(Eb / para.Nk, exact[:10], rtol=1e-10, atol=0)
