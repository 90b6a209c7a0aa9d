"""The call sequence of the Julia extension (the reference-side binding a maintainer adds),
replayed through ctypes.

The binding's create_gpu_linearmap receives a LinearMap, whose term list is the full list
(MomentumED/src/method/linear_map.jl:22-41), and its HilbertSubspace's bare state list
(the reference's CUDA extension uploads the same two things, MomentumEDCUDAExt.jl:94-113).
It recovers the upper triangle with `in >= out`, uploads it with upper_hermitian = 1, and
uploads the states with edm_basis_upload, without a system (or with one, when the caller
passes EDPara and the particle numbers). Both must give rank-table lookups and H x equal to
the oracle's restatement of linear_map.jl:100-118.
"""
import ctypes

import numpy as np
import pytest

import momentumed as me
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import check

pytestmark = pytest.mark.gpu


def _upload_op(ctx, bits, amp, out, inn, upper):
    amp = np.ascontiguousarray(amp, dtype=np.complex128)
    out = np.ascontiguousarray(out, dtype=np.uint64)
    inn = np.ascontiguousarray(inn, dtype=np.uint64)
    h = ctypes.c_void_p()
    check(ctx.L.edm_op_upload(ctx.h, bits, amp.ctypes.data, out.ctypes.data, inn.ctypes.data,
                              amp.size, int(upper), ctypes.byref(h)), ctx.h)
    return h


def _upload_basis(ctx, states, bits, sysp, sector):
    states = np.ascontiguousarray(states, dtype=np.uint64)
    h = ctypes.c_void_p()
    check(ctx.L.edm_basis_upload(ctx.h, states.ctypes.data, states.size, bits, sysp, sector,
                                 ctypes.byref(h)), ctx.h)
    return h


def _full_list(ctx, oph):
    n = ctypes.c_int64()
    check(ctx.L.edm_op_size(oph, ctypes.byref(n)), ctx.h)
    S = n.value
    amp = np.empty(S, np.complex128)
    out = np.empty(S, np.uint64)
    inn = np.empty(S, np.uint64)
    par = np.empty(S, np.uint64)
    check(ctx.L.edm_op_full_list(oph, amp.ctypes.data, out.ctypes.data, inn.ctypes.data,
                                 par.ctypes.data), ctx.h)
    return amp, out, inn, par


@pytest.mark.parametrize("name", ["small", "medium"])
def test_julia_binding_call_sequence(gpu_ctx, oracle_mod, name):
    ctx = gpu_ctx
    w = W.get(name)
    op = w.operator()
    bits = w.para.bits
    # the LinearMap's list: the full list, sorted (in, out)
    full = oracle_mod.full_list(oracle_mod.TermList(*op.arrays()), True)
    # the binding: up = [s for s in A.scat_list if s.in.n >= s.out.n]
    keep = full.inn >= full.out
    oph = _upload_op(ctx, bits, full.amp[keep], full.out[keep], full.inn[keep], True)
    try:
        a2, o2, i2, p2 = _full_list(ctx, oph)
        np.testing.assert_array_equal(o2, full.out)
        np.testing.assert_array_equal(i2, full.inn)
        np.testing.assert_array_equal(p2, full.parity)
        np.testing.assert_array_equal(a2, full.amp)
        sp = me.ED_momentum_subspaces(w.para, w.N_each, momentum_restriction=True,
                                      momentum_list=[w.sector])[0][0]
        states = sp.list
        sysd = dev.system_struct(w.para.k_list, w.para.Gk, w.para.Nc_mix, w.para.Nc_conserve,
                                 w.N_each)
        sector = me.preparation._momentum_list(w.para, w.N_each).index(w.sector)
        rng = np.random.default_rng(11)
        x = rng.normal(size=len(states)) + 1j * rng.normal(size=len(states))
        yo = oracle_mod.matvec(states, full, x)
        for sysp, sec, mode in ((None, -1, 2), (ctypes.byref(sysd), sector, 1)):
            bh = _upload_basis(ctx, states, bits, sysp, sec)
            try:
                m = ctypes.c_int32()
                check(ctx.L.edm_basis_rank_mode(bh, ctypes.byref(m)), ctx.h)
                assert m.value == mode
                y = np.empty_like(x)
                check(ctx.L.edm_matvec_host(ctx.h, oph, bh, x.ctypes.data, y.ctypes.data), ctx.h)
                assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
            finally:
                ctx.L.edm_basis_destroy(bh)
    finally:
        ctx.L.edm_op_destroy(oph)


def test_derived_rank_tables_on_a_list_that_is_not_closed(gpu_ctx, oracle_mod):
    """The first half of a sector splits into equal blocks (it gets derived tables) but the
    operator sends some of its states outside it: those targets must be dropped, as the
    binary search drops them, and no lookup may use the tables."""
    w = W.medium()
    op = w.operator()
    full = oracle_mod.full_list(oracle_mod.TermList(*op.arrays()), True)
    sp = me.ED_momentum_subspaces(w.para, w.N_each, momentum_restriction=True,
                                  momentum_list=[w.sector])[0][0]
    half = sp.list[: len(sp.list) // 2].copy()
    b = dev.DeviceBasis.upload(gpu_ctx, half, w.para.bits)
    assert b.rank_mode == 2
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    x = np.random.default_rng(3).normal(size=len(half)) + 0j
    y = dev.ROCmLinearMap(dop, b)(x)
    yo = oracle_mod.matvec(half, full, x)
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    b.drop_rank()
    np.testing.assert_array_equal(dev.ROCmLinearMap(dop, b)(x), y)
