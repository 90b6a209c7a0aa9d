"""Eigenvector consumers on the device: brackets, the one-body density matrix, and the
kshift step overlaps of many-body Berry connections.

Mirrors EDCore/src/multiplication.jl:153-307 (ED_bracket, ED_bracket_threaded, one- and
two-operator forms),
MomentumED/src/analysis/particle_reduced_density_matrix.jl:4-88 (PES_1rdm,
PES_MomtBlock_rdm) and
MomentumED/src/analysis/manybody_quantum_geometry.jl:1-91 (ED_connection_gaugefixing!,
ED_step_inner_prod). The vectors are MBS64Vector (host arrays bound to a subspace); each call
copies them to the device, runs the libedm.so kernel and returns host results.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check
from .core import HilbertSubspace, MBOperator, MBS64Vector
from .device import DeviceOperator, device_basis_of, get_context

__all__ = ["ED_bracket", "ED_bracket_threaded", "PES_1rdm", "PES_MomtBlock_rdm",
           "ED_step_inner_prod", "ED_connection_gaugefixing"]


def _dev(vec: np.ndarray, ctx):
    import torch
    return torch.from_numpy(np.ascontiguousarray(vec, dtype=np.complex128)).to(
        f"cuda:{ctx.device}")


def ED_bracket(bra: MBS64Vector, op: MBOperator, *rest):
    """<bra|op|ket> (multiplication.jl:270-280), computed as bra^H (O ket) with the
    operator's full list (an upper-Hermitian operator includes the adjoints of its
    off-diagonal terms). bra and ket may live in different subspaces (then O maps the ket's
    subspace into the bra's; the reference assumes one subspace for upper-Hermitian
    operators). ED_bracket(bra, op2, op1, ket) (multiplication.jl:281-291) is
    <bra|op2 op1|ket>: both full lists applied from the left to each ket state on the device,
    final states looked up in the bra's subspace (edm_bracket2)."""
    if len(rest) == 2:
        return _bracket2(bra, op, rest[0], rest[1])
    if len(rest) != 1:
        raise TypeError("ED_bracket(bra, op, ket) or ED_bracket(bra, op2, op1, ket)")
    ket = rest[0]
    if bra.space.bits != ket.space.bits:
        raise AssertionError("bra and ket have different bits")
    import torch
    ctx = get_context(0)
    bb = device_basis_of(bra.space, ctx)
    kb = bb if ket.space is bra.space else device_basis_of(ket.space, ctx)
    amp, out, inn = op.arrays()
    dop = DeviceOperator(ctx, bra.space.bits, amp, out, inn, op.upper_hermitian)
    x = _dev(ket.vec, ctx)
    y = torch.empty(len(bra.vec), dtype=torch.complex128, device=x.device)
    torch.cuda.current_stream(x.device).synchronize()
    check(ctx.L.edm_matvec_between(ctx.h, dop.h, bb.h, kb.h, x.data_ptr(), y.data_ptr()), ctx.h)
    ctx.synchronize()
    return complex(torch.vdot(_dev(bra.vec, ctx), y).item())


def _bracket2(bra: MBS64Vector, op2: MBOperator, op1: MBOperator, ket: MBS64Vector) -> complex:
    if bra.space.bits != ket.space.bits:
        raise AssertionError("bra and ket have different bits")
    import torch
    ctx = get_context(0)
    bb = device_basis_of(bra.space, ctx)
    kb = bb if ket.space is bra.space else device_basis_of(ket.space, ctx)
    ops = []
    for op in (op2, op1):
        amp, out, inn = op.arrays()
        ops.append(DeviceOperator(ctx, bra.space.bits, amp, out, inn, op.upper_hermitian))
    xb = _dev(bra.vec, ctx)
    xk = _dev(ket.vec, ctx)
    torch.cuda.current_stream(xb.device).synchronize()
    res = np.zeros(2)
    check(ctx.L.edm_bracket2(ctx.h, ops[0].h, ops[1].h, bb.h, kb.h, xb.data_ptr(),
                             xk.data_ptr(), res.ctypes.data), ctx.h)
    return complex(res[0], res[1])


ED_bracket_threaded = ED_bracket


def PES_1rdm(psi: MBS64Vector) -> np.ndarray:
    """One-body reduced density matrix rho[i, j] = <psi|c_i^+ c_j|psi> (bits x bits,
    0-based here for orbitals 1..bits of the reference), computed for i <= j on the device
    and completed by rho[j, i] = conj(rho[i, j]) (particle_reduced_density_matrix.jl:4-33)."""
    bits = psi.space.bits
    ctx = get_context(0)
    b = device_basis_of(psi.space, ctx)
    pairs = np.array([(i, j) for i in range(bits) for j in range(bits) if i <= j],
                     dtype=np.int32).reshape(-1)
    n = pairs.size // 2
    out = np.zeros(2 * n, dtype=np.float64)
    x = _dev(psi.vec, ctx)
    import torch
    torch.cuda.current_stream(x.device).synchronize()
    check(ctx.L.edm_rdm1(ctx.h, b.h, x.data_ptr(), n, pairs.ctypes.data, out.ctypes.data), ctx.h)
    vals = out[0::2] + 1j * out[1::2]
    rdm = np.zeros((bits, bits), dtype=np.complex128)
    for p in range(n):
        i, j = pairs[2 * p], pairs[2 * p + 1]
        rdm[i, j] = vals[p]
    for i in range(bits):
        for j in range(i):
            rdm[i, j] = np.conj(rdm[j, i])
    return rdm


def PES_MomtBlock_rdm(para, psi: MBS64Vector, part_subspace, ph_transform=None) -> np.ndarray:
    """Particle reduced density matrix of psi on one momentum block of part A
    (particle_reduced_density_matrix.jl:37-88), |A| x |A| Hermitian. Computed on the device
    as M M^H: M[a, b] = (-1)^occ(s, a) psi[s] for s ^ ph = a | b with b in the complementary
    block, filled by edm_pes_block_matrix, and the product as one dense GEMM on the GPU.

    ph_transform (one flag per conserved component, default all false) particle-hole
    transforms the flagged components before the split, as the reference does: part A's
    states and the complement are then occupation patterns of the transformed state, and the
    signs are taken from the untransformed states. The complement is the block of the
    remaining particles of every component of the transformed state and the remaining
    momentum."""
    import torch
    from .preparation import ED_momentum_subspaces, MBS_totalmomentum, momentum_residue
    if ph_transform is None:
        ph_transform = [False] * para.Nc_conserve
    if len(ph_transform) != para.Nc_conserve:
        raise AssertionError("lengths of subspaces and particle_hole should == "
                             "para.Nc_conserve.")
    if psi.space.bits != para.Nk * para.Nc:
        raise AssertionError("bits == para.Nk * para.Nc")
    width = para.Nk * para.Nc_mix
    comp = (1 << width) - 1
    ph = 0
    for c, flag in enumerate(ph_transform):
        if flag:
            ph |= comp << (c * width)
    A = np.ascontiguousarray(part_subspace.list, dtype=np.uint64)
    nA = len(A)
    rdm = np.zeros((nA, nA), dtype=np.complex128)
    if nA == 0 or len(psi.vec) == 0:
        return rdm

    def numbers(state):
        return tuple(bin((int(state) >> (c * width)) & comp).count("1")
                     for c in range(para.Nc_conserve))
    s0 = int(psi.space.list[0]) ^ ph        # every state of a sector gives the same numbers
    n_B = tuple(a - b for a, b in zip(numbers(s0), numbers(A[0])))
    if any(v < 0 for v in n_B):
        return rdm
    k_s, k_A = MBS_totalmomentum(para, s0), MBS_totalmomentum(para, int(A[0]))
    k_B = momentum_residue(tuple(a - b for a, b in zip(k_s, k_A)), para.Gk)
    ctx = get_context(0)
    pb = device_basis_of(psi.space, ctx)
    x = _dev(psi.vec, ctx)
    acc = torch.zeros((nA, nA), dtype=torch.complex128, device=x.device)
    counts = np.zeros(2, dtype=np.int64)
    found = 0

    def add_block(B):
        nonlocal found
        M = torch.zeros((nA, max(B.length, 1)), dtype=torch.complex128, device=x.device)
        torch.cuda.current_stream(x.device).synchronize()
        check(ctx.L.edm_pes_block_matrix(ctx.h, pb.h, x.data_ptr(), nA, A.ctypes.data, ph,
                                         device_basis_of(B, ctx).h, M.data_ptr(),
                                         counts.ctypes.data), ctx.h)
        found += int(counts[1])
        if counts[1]:
            acc.add_(M @ M.conj().T)
    # the complement block of the first state; with particle-hole transformed components
    # and several conserved components the complements of one sector spread over several
    # momenta, and the other blocks are visited as well
    if sum(n_B) == 0:   # part A holds every particle: the complement is the empty state
        add_block(HilbertSubspace([0], bits=psi.space.bits))
    else:
        subs, ks = ED_momentum_subspaces(para, n_B, momentum_restriction=True,
                                         momentum_list=[k_B])
        # an empty block still counts the (state, part) pairs
        add_block(subs[0] if subs else HilbertSubspace([], bits=psi.space.bits))
        if found < counts[0]:
            subs, ks = ED_momentum_subspaces(para, n_B)
            for sub, k in zip(subs, ks):
                if k != k_B:
                    add_block(sub)
    if found != counts[0]:
        raise ValueError("psi's states do not share the particle numbers of part A plus one "
                         "complement (psi must live in one particle-number sector)")
    ctx.synchronize()
    return acc.cpu().numpy()


def ED_step_inner_prod(psi_f, psi_i, orbital_inner_prod):
    """Many-body connection step: sum_x conj(f[x]) i[x] prod_{k occupied in x} u[k]; for lists
    of g vectors each, the g x g matrix with entry [a, b] from f_a and i_b
    (manybody_quantum_geometry.jl:43-91)."""
    single = isinstance(psi_f, MBS64Vector)
    fl = [psi_f] if single else list(psi_f)
    il = [psi_i] if single else list(psi_i)
    if len(fl) != len(il):
        raise AssertionError("Number of eigenvectors must match in non-Abelian connection")
    space = fl[0].space
    for v in fl + il:
        if not (v.space is space or np.array_equal(v.space.list, space.list)):
            raise AssertionError("The Hilbert subspaces of eigenvectors must match.")
    u = np.ascontiguousarray(np.asarray(orbital_inner_prod, dtype=np.complex128).reshape(-1))
    if u.size != space.bits:
        raise AssertionError(f"orbital inner product length should be {space.bits}.")
    g = len(fl)
    ctx = get_context(0)
    b = device_basis_of(space, ctx)
    df = [_dev(v.vec, ctx) for v in fl]
    di = [_dev(v.vec, ctx) for v in il]
    import torch
    torch.cuda.current_stream(df[0].device).synchronize()
    pf = (ctypes.c_void_p * g)(*[t.data_ptr() for t in df])
    pi = (ctypes.c_void_p * g)(*[t.data_ptr() for t in di])
    out = np.zeros(2 * g * g, dtype=np.float64)
    check(ctx.L.edm_step_inner_prod(ctx.h, b.h, g, pf, pi, u.ctypes.data, out.ctypes.data), ctx.h)
    m = (out[0::2] + 1j * out[1::2]).reshape(g, g)
    return complex(m[0, 0]) if single else m


def ED_connection_gaugefixing(psi, suggest_dims=None, warning_tol: float = 1e-8):
    """Fix the overall phase of eigenvectors in place before a Wilson-loop integral
    (manybody_quantum_geometry.jl:14-40): one vector is rotated so that its element at
    position suggest_dims (1-based) is real positive; for a list, the first vector is rotated
    by the phase of det[psi_j[suggest_dims[i]]]. Host arithmetic on host vectors."""
    import warnings
    if isinstance(psi, MBS64Vector):
        pos = 1 if suggest_dims is None else int(suggest_dims)
        if not 1 <= pos <= len(psi.vec):
            raise AssertionError("suggested vector locates outside the Hilbert space.")
        x = psi.vec[pos - 1]
        if abs(x) < warning_tol:
            warnings.warn(f"Gauge fixing may be unstable. Element at position {pos} has "
                          f"amplitude {abs(x)}.")
        psi.vec *= np.exp(-1j * np.angle(x))
        return psi
    vecs = list(psi)
    dims = list(range(1, len(vecs) + 1)) if suggest_dims is None else [int(d) for d in suggest_dims]
    if len(dims) != len(vecs):
        raise AssertionError("number of suggested vectors mismatches number of eigenvectors.")
    mat = np.array([[v.vec[i - 1] for v in vecs] for i in dims])
    x = np.linalg.det(mat)
    if abs(x) < warning_tol:
        warnings.warn(f"Gauge fixing may be unstable. Determinant of elements at position "
                      f"{dims} has amplitude {abs(x)}.")
    vecs[0].vec *= np.exp(-1j * np.angle(x))
    return vecs

