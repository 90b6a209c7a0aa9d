"""The stored-matrix (:sparse) path on the device: CSR build, SpMV, Krylov solve.

Bars: the CSR structure (row offsets, ascending column indices) equals the oracle's
connectivity with equal (row, column) entries merged, bit for bit; values within 1e-15 of
max |H_ij| (duplicates are summed in a different order); y = A x within 1e-13 of max |y| of
the oracle and of the matrix-free map; eigenvalues within 1e-10 relative of dense LAPACK /
ARPACK and of the map method.
"""
import numpy as np
import pytest
import scipy.sparse as sps

import momentumed as me
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import EDMError

pytestmark = pytest.mark.gpu


def _system(para, N_each):
    return dev.system_struct(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, N_each)


def _oracle_full(O, op):
    amp, out, inn = op.arrays()
    return O.full_list(O.TermList(amp, out, inn), op.upper_hermitian)


def _oracle_csr(O, basis, full):
    rp, cols, terms, negs = O.connectivity(basis, full)
    vals = full.amp[terms] * np.where(negs, -1.0, 1.0)
    H = sps.csr_matrix((vals, cols, rp), shape=(len(basis), len(basis)))
    H.sum_duplicates()
    H.sort_indices()
    return H


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _compare(m, Href):
    rp, cols, vals = m.download()
    np.testing.assert_array_equal(rp, Href.indptr)
    np.testing.assert_array_equal(cols, Href.indices)
    scale = np.abs(Href.data).max()
    assert np.abs(vals - Href.data).max() <= 1e-15 * scale * 4


def test_csr_small_all_sectors(gpu_ctx, oracle_mod):
    print("[csr] small 3x5 N=5, all sectors", flush=True)
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    subspaces, _ = me.ED_momentum_subspaces(w.para, 5)
    for k, sp in enumerate(subspaces):
        m = dev.create_gpu_matrix(op, sp)
        Href = _oracle_csr(oracle_mod, sp.list, full)
        _compare(m, Href)
        x = _rand(sp.length, k)
        y = m(x)
        yo = oracle_mod.matvec(sp.list, full, x)
        assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()


def test_csr_medium_sector_against_oracle_and_map(gpu_ctx, oracle_mod):
    print("[csr] medium 3x9 N=9 sector (0,0)", flush=True)
    w = W.medium()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 9, momentum_restriction=True,
                                  momentum_list=[(0, 0)])[0][0]
    m = dev.create_gpu_matrix(op, sp)
    full = _oracle_full(oracle_mod, op)
    _compare(m, _oracle_csr(oracle_mod, sp.list, full))
    x = _rand(sp.length, 5)
    H = dev.create_gpu_linearmap(op, sp)
    ym = H(x)
    ys = m(x)
    assert np.abs(ys - ym).max() <= 1e-13 * np.abs(ym).max()


def test_csr_partial_basis_and_two_components(gpu_ctx, oracle_mod):
    """Binary-search partner lookup (targets outside the basis dropped) and a two-component
    sector with rank tables."""
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    lists, _ = oracle_mod.enumerate_sectors(w.para.k_list, w.para.Gk, 1, 1, [5])
    part = lists[0][::3].copy()
    b = dev.DeviceBasis.upload(gpu_ctx, part, w.para.bits, _system(w.para, w.N_each), 0)
    assert b.rank_mode == 2   # derived tables; the closure check sends lookups to search
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    _compare(dev.DeviceCSR(dop, b), _oracle_csr(oracle_mod, part, full))
    para = W.batched_para(0.7)
    op2 = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sp2 = me.ED_momentum_subspaces(para, (3, 3))[0][4]
    _compare(dev.create_gpu_matrix(op2, sp2),
             _oracle_csr(oracle_mod, sp2.list, _oracle_full(oracle_mod, op2)))


def test_edsolve_sparse_against_lapack_and_map(gpu_ctx, oracle_mod):
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    subspaces, _ = me.ED_momentum_subspaces(w.para, 5)
    for sp in subspaces[:5]:
        E, V = me.EDsolve(sp, op, N=w.nev, method="sparse", device="gpu", method_info=False)
        Em, _ = me.EDsolve(sp, op, N=w.nev, method="map", device="gpu", method_info=False)
        ev, evec = np.linalg.eigh(oracle_mod.dense_matrix(sp.list, full))
        np.testing.assert_allclose(E, ev[: w.nev], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(E, Em, rtol=1e-10, atol=1e-12)
        Hd = oracle_mod.dense_matrix(sp.list, full)
        for i, v in enumerate(V):
            assert np.linalg.norm(Hd @ v.vec - E[i] * v.vec) < 1e-9


def test_edsolve_float32_results(gpu_ctx, oracle_mod):
    """element_type=Float32 (MomentumED.jl:220,378): the sparse and dense methods return
    Float32 energies and ComplexF32 vectors, the Float64 results rounded (the device computes
    in complex FP64); the map ignores element_type, as the reference's map does."""
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    sp = me.ED_momentum_subspaces(w.para, 5)[0][0]
    ev = np.linalg.eigvalsh(oracle_mod.dense_matrix(sp.list, full))[: w.nev]
    for method in ("sparse", "dense"):
        E64, V64 = me.EDsolve(sp, op, N=w.nev, method=method, method_info=False)
        E32, V32 = me.EDsolve(sp, op, N=w.nev, method=method, element_type=np.float32,
                              method_info=False)
        assert E64.dtype == np.float64 and E32.dtype == np.float32, method
        assert all(v.vec.dtype == np.complex64 for v in V32), method
        np.testing.assert_array_equal(E32, E64.astype(np.float32))
        for a, b in zip(V32, V64):
            np.testing.assert_array_equal(a.vec, b.vec.astype(np.complex64))
        np.testing.assert_allclose(E32, ev, rtol=1e-6, atol=1e-6)
        E32s, _ = me.EDsolve(sp, op, N=w.nev, method=method, element_type="Float32",
                             method_info=False)
        np.testing.assert_array_equal(E32s, E32)
    Em, Vm = me.EDsolve(sp, op, N=w.nev, method="map", element_type=np.float32,
                        method_info=False)
    assert Em.dtype == np.float64 and Vm[0].vec.dtype == np.complex128
    np.testing.assert_allclose(Em, ev, rtol=1e-10, atol=1e-12)


def test_edsolve_sparse_medium_against_arpack(gpu_ctx, oracle_mod):
    import scipy.sparse.linalg as spl
    w = W.medium()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 9, momentum_restriction=True,
                                  momentum_list=[(0, 0)])[0][0]
    E, V = me.EDsolve(sp, op, N=w.nev, method="sparse", device="gpu", method_info=False,
                      krylovdim=w.krylovdim, tol=w.tol)
    Hs = _oracle_csr(oracle_mod, sp.list, _oracle_full(oracle_mod, op))
    ref = np.sort(spl.eigsh(Hs, k=w.nev, which="SA", tol=1e-14, ncv=60)[0])
    np.testing.assert_allclose(E, ref, rtol=1e-10, atol=0)


def test_edsolve_scatter_list_form(gpu_ctx):
    """EDsolve(subspace, scat_list1, scat_list2; ...) (MomentumED.jl:217-231) equals the
    MBOperator form."""
    w = W.small()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][2]
    l1 = me.ED_scatterlist_onebody(w.para)
    l2 = me.ED_scatterlist_twobody(w.para)
    E1, _ = me.EDsolve(sp, l1, l2, N=4, method="sparse", method_info=False)
    E2, _ = me.EDsolve(sp, me.MBOperator(l1, l2, upper_hermitian=True), N=4, method="sparse",
                       method_info=False)
    np.testing.assert_array_equal(E1, E2)


def test_csr_rejects_generic_terms(gpu_ctx):
    bits = 8
    t = [me.Scatter(1.0, 1, 2, 3, 4, 5, 6, bits=bits, upper_hermitian=True)]
    op = me.MBOperator(t, upper_hermitian=True)
    sp = me.HilbertSubspace([s for s in range(1 << bits) if bin(s).count("1") == 3], bits=bits)
    with pytest.raises(EDMError):
        dev.create_gpu_matrix(op, sp)


def test_edsolve_dense_and_method_switching(gpu_ctx, oracle_mod, caplog):
    """method="dense" (MomentumED.jl:340-369) on the device against LAPACK, and the
    reference's automatic switches (MomentumED.jl:233-244): sparse below min_sparse_dim runs
    dense, dense above max_dense_dim runs sparse, scatter lists allow dense up to 500."""
    import logging
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    sp = me.ED_momentum_subspaces(w.para, 5)[0][3]
    Hd = oracle_mod.dense_matrix(sp.list, full)
    ev = np.linalg.eigvalsh(Hd)
    E, V = me.EDsolve(sp, op, N=w.nev, method="dense", max_dense_dim=500, method_info=False)
    np.testing.assert_allclose(E, ev[: w.nev], rtol=0, atol=1e-12 * np.abs(ev).max())
    for i, v in enumerate(V):
        assert np.linalg.norm(Hd @ v.vec - E[i] * v.vec) < 1e-10
    caplog.set_level(logging.INFO, logger="momentumed")
    caplog.clear()
    E2, _ = me.EDsolve(sp, op, N=w.nev, min_sparse_dim=300, max_dense_dim=500)
    assert "switch to method=:dense" in caplog.text
    np.testing.assert_allclose(E2, E, rtol=0, atol=1e-12 * np.abs(ev).max())
    caplog.clear()
    assert sp.length == 200
    E3, _ = me.EDsolve(sp, op, N=w.nev, method="dense", max_dense_dim=150)
    assert "switch to method=:sparse" in caplog.text
    np.testing.assert_allclose(E3, ev[: w.nev], rtol=1e-10, atol=1e-12)
    caplog.clear()
    l2 = me.ED_scatterlist_twobody(w.para)
    l1 = me.ED_scatterlist_onebody(w.para)
    big = me.ED_momentum_subspaces(w.para, 5)[0][0]
    assert big.length == 201
    E5, _ = me.EDsolve(big, op, N=w.nev, method="dense")       # 201 > 200 for an MBOperator
    assert "switch to method=:sparse" in caplog.text
    caplog.clear()
    E4, _ = me.EDsolve(sp, l1, l2, N=w.nev, method="dense")    # scatter lists: up to 500
    assert "switch to method" not in caplog.text
    np.testing.assert_allclose(E4, E, rtol=0, atol=1e-12 * np.abs(ev).max())


def test_reference_constructor_forms(gpu_ctx, capsys):
    """create_gpu_matrix(SparseHmltMatrix(subspace, scats), :rocm) and
    create_gpu_linearmap(LinearMap(H, subspace), :rocm) as reference scripts write them, with
    GPU_MEMORY_MONITOR reporting the context's HBM use after each solve."""
    w = W.small()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][2]
    scats = me.ED_scatterlist_onebody(w.para) + me.ED_scatterlist_twobody(w.para)
    H = me.SparseHmltMatrix(sp, scats)
    assert H.size() == (sp.length, sp.length)
    me.GPU_MEMORY_MONITOR["value"] = True
    try:
        v1, _, i1 = me.krylov_matrix_solve(me.create_gpu_matrix(H, "rocm"), 4)
        op = me.MBOperator(scats, upper_hermitian=True)
        v2, _, _ = me.krylov_map_solve(me.create_gpu_linearmap(me.LinearMap(op, sp), "rocm"), 4)
    finally:
        me.GPU_MEMORY_MONITOR["value"] = False
    assert "GiB in use" in capsys.readouterr().out
    E, _ = me.EDsolve(sp, op, N=4, method_info=False)
    np.testing.assert_allclose(v1, E, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(v2, E, rtol=1e-10, atol=1e-12)
    assert i1["converged"] >= 4
