"""edm_bracket2_noalloc, the allocation-free form of edm_bracket2, through the C-ABI.

It runs the same kernels in the same order with its partial sums in the library's file-scope
scratch array, so its result must equal edm_bracket2's bit for bit: checked on two calls in a
row, the second with a smaller input (fewer chunk partials than the first left behind), and
on an empty ket basis.
"""
import ctypes

import numpy as np
import pytest

import momentumed as me
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import check

pytestmark = pytest.mark.gpu


def _density_op(q, f, nx, ny):
    """sum_k f[k] c^+_{k+q} c_k on an nx x ny mesh (orbital = kx + nx*ky + 1)."""
    terms = []
    for kx in range(nx):
        for ky in range(ny):
            a = (kx + q[0]) % nx + nx * ((ky + q[1]) % ny) + 1
            b = kx + nx * ky + 1
            terms.append(me.Scatter(complex(f[b - 1]), a, b, bits=nx * ny))
    return me.MBOperator(terms, upper_hermitian=False)


def _both(ctx, op2, op1, bra_sp, ket_sp, bra, ket):
    import torch
    bb = dev.device_basis_of(bra_sp, ctx)
    kb = dev.device_basis_of(ket_sp, ctx)
    d2 = dev.DeviceOperator.from_mboperator(ctx, op2, bra_sp.bits)
    d1 = dev.DeviceOperator.from_mboperator(ctx, op1, bra_sp.bits)
    xb = torch.from_numpy(bra).cuda() if bra.size else torch.zeros(1, dtype=torch.complex128,
                                                                   device="cuda:0")
    xk = torch.from_numpy(ket).cuda() if ket.size else torch.zeros(1, dtype=torch.complex128,
                                                                   device="cuda:0")
    torch.cuda.synchronize()
    ref = np.zeros(2)
    check(ctx.L.edm_bracket2(ctx.h, d2.h, d1.h, bb.h, kb.h, xb.data_ptr(), xk.data_ptr(),
                             ref.ctypes.data), ctx.h)
    out = torch.full((1,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    check(ctx.L.edm_bracket2_noalloc(ctx.h, d2.h, d1.h, bb.h, kb.h, xb.data_ptr(),
                                     xk.data_ptr(), out.data_ptr()), ctx.h)
    ctx.synchronize()
    got = out.cpu().numpy()[0]
    return complex(ref[0], ref[1]), complex(got)


def test_bracket2_noalloc_equals_bracket2(gpu_ctx):
    rng = np.random.default_rng(5)
    # call 1: rho(-q) rho(q) on the 3x9 N=9 sector (0,0), 173,613 states (679 chunk partials)
    w = W.medium()
    sp = me.ED_momentum_subspaces(w.para, 9, momentum_restriction=True,
                                  momentum_list=[(0, 0)])[0][0]
    f = rng.normal(size=27) + 1j * rng.normal(size=27)
    op1 = _density_op((1, 2), f, 3, 9)
    op2 = _density_op((2, 7), f.conj(), 3, 9)
    v1 = rng.normal(size=sp.length) + 1j * rng.normal(size=sp.length)
    v2 = rng.normal(size=sp.length) + 1j * rng.normal(size=sp.length)
    ref, got = _both(gpu_ctx, op2, op1, sp, sp, v1, v2)
    print(f"[noalloc] medium: {ref} / {got}", flush=True)
    assert ref != 0 and got == ref
    # call 2, smaller: H H on the 3x5 N=5 sector (0,0), 201 states (one chunk partial)
    ws = W.small()
    sps = me.ED_momentum_subspaces(ws.para, 5)[0][0]
    H = ws.operator()
    a = rng.normal(size=sps.length) + 1j * rng.normal(size=sps.length)
    b = rng.normal(size=sps.length) + 1j * rng.normal(size=sps.length)
    ref, got = _both(gpu_ctx, H, H, sps, sps, a, b)
    print(f"[noalloc] small: {ref} / {got}", flush=True)
    assert ref != 0 and got == ref
    # <v|H H|v> = ||H v||^2 is real and positive
    ref, got = _both(gpu_ctx, H, H, sps, sps, a, a)
    assert got == ref and got.real > 0 and abs(got.imag) <= 1e-12 * got.real
    # an empty ket basis gives 0
    empty = me.HilbertSubspace(np.zeros(0, dtype=np.uint64), bits=15)
    ref, got = _both(gpu_ctx, H, H, sps, empty, a, np.zeros(0, complex))
    assert ref == 0 and got == 0


def test_bracket2_noalloc_rejects_bad_arguments(gpu_ctx):
    ws = W.small()
    sps = me.ED_momentum_subspaces(ws.para, 5)[0][0]
    bb = dev.device_basis_of(sps, gpu_ctx)
    d = dev.DeviceOperator.from_mboperator(gpu_ctx, ws.operator(), sps.bits)
    rc = gpu_ctx.L.edm_bracket2_noalloc(gpu_ctx.h, d.h, d.h, bb.h, bb.h, None, None, None)
    assert rc != 0
