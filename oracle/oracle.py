"""CPU oracle for the momentum-sector ED hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker (or, for cpu_baseline, as the timed CPU algorithm). The
product (the ``momentumed`` package and ``libedm.so``) never imports or calls it.

Contents
  * ctypes bindings to ``liboracle.so`` (edm_oracle.c): parity mask, apply, full-list
    expansion, basis enumeration, row-gather matvec (Hx), connectivity.
  * ``scatterlist_twobody`` / ``scatterlist_onebody``: restatement of the reference's term
    enumeration (MomentumED/src/preparation/scat_list.jl:78-142,425-572) with the Scatter
    constructor rules (EDCore/src/types/scattering.jl:63-153) and ``sort_merge_scatlist``
    (scattering.jl:258-282).
  * ``jw_hamiltonian``: an independent dense Hamiltonian by Jordan-Wigner second
    quantisation (textbook), used to pin the term builder and the apply/sign rules.

Pins of each piece are listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import itertools
import os
from fractions import Fraction
from typing import Callable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib() -> ctypes.CDLL:
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        u64, i64, i32, dbl = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        vp = ctypes.c_void_p
        L.orc_parity_mask.argtypes = [u64]
        L.orc_parity_mask.restype = u64
        L.orc_apply_right.argtypes = [u64, u64, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(i32)]
        L.orc_apply_right.restype = i32
        L.orc_apply_left.argtypes = L.orc_apply_right.argtypes
        L.orc_apply_left.restype = i32
        L.orc_search.argtypes = [vp, i64, u64]
        L.orc_search.restype = i64
        L.orc_full_list.argtypes = [i64, vp, vp, vp, i32, vp, vp, vp, vp]
        L.orc_full_list.restype = i64
        L.orc_matvec_rows.argtypes = [i64, vp, i64, vp, vp, vp, vp, vp, vp, i64, i64, i32]
        L.orc_matvec_rows.restype = None
        L.orc_matvec_rowlist.argtypes = [i64, vp, i64, vp, vp, vp, vp, vp, vp, vp, i64, i32]
        L.orc_matvec_rowlist.restype = None
        L.orc_connectivity.argtypes = [i64, vp, i64, vp, vp, vp, vp, i64, i64, vp, vp, vp, vp, vp, i32]
        L.orc_connectivity.restype = None
        L.orc_enumerate.argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp]
        L.orc_enumerate.restype = i64
        L.orc_num_sectors.argtypes = [i32, i32, i32, vp, vp, vp]
        L.orc_num_sectors.restype = i64
        L.orc_bracket2.argtypes = [i64, vp, vp, i64, vp, vp, i64, vp, vp, vp, vp, i64, vp, vp,
                                   vp, vp, vp]
        L.orc_bracket2.restype = None
        L.orc_pes_block_rdm.argtypes = [i64, vp, vp, i64, vp, u64, vp]
        L.orc_pes_block_rdm.restype = None
        _ = dbl
        _LIB = L
    return _LIB


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


# ----------------------------------------------------------------------------- bit rules
def parity_mask(occ: int) -> int:
    return int(lib().orc_parity_mask(occ))


def apply_right(state: int, out: int, inn: int, parity: int):
    t = ctypes.c_uint64()
    n = ctypes.c_int()
    ok = lib().orc_apply_right(state, out, inn, parity, ctypes.byref(t), ctypes.byref(n))
    return (bool(ok), int(t.value), int(n.value))


def apply_left(state: int, out: int, inn: int, parity: int):
    t = ctypes.c_uint64()
    n = ctypes.c_int()
    ok = lib().orc_apply_left(state, out, inn, parity, ctypes.byref(t), ctypes.byref(n))
    return (bool(ok), int(t.value), int(n.value))


# --------------------------------------------------------------------------- full list
class TermList:
    """Structure-of-arrays term list: amp (complex128), out/in/parity (uint64)."""

    def __init__(self, amp, out, inn, parity=None):
        self.amp = np.ascontiguousarray(amp, dtype=np.complex128)
        self.out = np.ascontiguousarray(out, dtype=np.uint64)
        self.inn = np.ascontiguousarray(inn, dtype=np.uint64)
        if parity is None:
            parity = np.array([parity_mask(int(o)) ^ parity_mask(int(i))
                               for o, i in zip(self.out, self.inn)], dtype=np.uint64)
        self.parity = np.ascontiguousarray(parity, dtype=np.uint64)

    def __len__(self):
        return len(self.amp)


def full_list(terms: TermList, upper_hermitian: bool) -> TermList:
    """MomentumED/src/method/linear_map.jl:22-41."""
    S = len(terms)
    amp_f = np.zeros(2 * S + 1, dtype=np.complex128)
    out_f = np.zeros(2 * S + 1, dtype=np.uint64)
    in_f = np.zeros(2 * S + 1, dtype=np.uint64)
    par_f = np.zeros(2 * S + 1, dtype=np.uint64)
    n = lib().orc_full_list(S, _p(terms.amp), _p(terms.out), _p(terms.inn), int(upper_hermitian),
                            _p(amp_f), _p(out_f), _p(in_f), _p(par_f))
    return TermList(amp_f[:n].copy(), out_f[:n].copy(), in_f[:n].copy(), par_f[:n].copy())


# --------------------------------------------------------------------------- enumeration
def enumerate_sectors(k_list: np.ndarray, Gk: Sequence[int], Nc_mix: int, Nc_conserve: int,
                      N_each: Sequence[int]):
    """momentum_decomposition.jl:188-262. Returns (list of uint64 arrays, list of momenta),
    empty sectors dropped (:257-259)."""
    k_list = np.asarray(k_list, dtype=np.int64)
    dim, nk = k_list.shape
    kl = np.ascontiguousarray(k_list.T.reshape(-1))   # column-major: k_list[d + dim*k]
    gk = np.ascontiguousarray(np.asarray(Gk, dtype=np.int64))
    ne = np.ascontiguousarray(np.asarray(N_each, dtype=np.int64))
    L = lib()
    nsec = L.orc_num_sectors(dim, nk, Nc_conserve, _p(kl), _p(gk), _p(ne))
    counts = np.zeros(nsec, dtype=np.int64)
    r = L.orc_enumerate(dim, nk, Nc_mix, Nc_conserve, _p(kl), _p(gk), _p(ne), _p(counts), 0, 0)
    if r < 0:
        raise ValueError("bad enumeration input")
    offsets = np.zeros(nsec, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)[:-1]
    states = np.zeros(int(counts.sum()) + 1, dtype=np.uint64)
    L.orc_enumerate(dim, nk, Nc_mix, Nc_conserve, _p(kl), _p(gk), _p(ne), _p(counts),
                    _p(offsets), _p(states))
    ntot = int(ne.sum())
    ranges = []
    for d in range(dim):
        if Gk[d] != 0:
            ranges.append(range(0, Gk[d]))
        else:
            ranges.append(range(int(k_list[d].min()) * ntot, int(k_list[d].max()) * ntot + 1))
    moms = []
    idx = np.indices([len(r) for r in ranges]).reshape(dim, -1)
    # first coordinate fastest (Iterators.product order)
    flat = np.ravel_multi_index(idx, [len(r) for r in ranges], order="F")
    order = np.argsort(flat)
    for j in order:
        moms.append(tuple(int(ranges[d][idx[d, j]]) for d in range(dim)))
    lists, ks = [], []
    for s in range(nsec):
        if counts[s]:
            lists.append(states[offsets[s]:offsets[s] + counts[s]].copy())
            ks.append(moms[s])
    return lists, ks


# --------------------------------------------------------------------------- matvec
def matvec(basis: np.ndarray, full: TermList, x: np.ndarray, rows=None, nthreads: int = 0):
    """linear_map.jl:100-118 (Hx). ``rows``: None (all), (row0, nrows) or an index array."""
    basis = np.ascontiguousarray(basis, dtype=np.uint64)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    D = len(basis)
    L = lib()
    if rows is None or isinstance(rows, tuple):
        row0, nrows = (0, D) if rows is None else rows
        y = np.zeros(nrows, dtype=np.complex128)
        L.orc_matvec_rows(D, _p(basis), len(full), _p(full.amp), _p(full.out), _p(full.inn),
                          _p(full.parity), _p(x), _p(y), row0, nrows, nthreads)
    else:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        y = np.zeros(len(rows), dtype=np.complex128)
        L.orc_matvec_rowlist(D, _p(basis), len(full), _p(full.amp), _p(full.out), _p(full.inn),
                             _p(full.parity), _p(x), _p(y), _p(rows), len(rows), nthreads)
    return y


def connectivity(basis: np.ndarray, full: TermList, row0: int = 0, nrows: int | None = None,
                 nthreads: int = 0):
    """Per-row (col, term, neg) lists in full-list order. Returns rowptr, cols, terms, negs."""
    basis = np.ascontiguousarray(basis, dtype=np.uint64)
    D = len(basis)
    if nrows is None:
        nrows = D - row0
    L = lib()
    counts = np.zeros(nrows, dtype=np.int64)
    L.orc_connectivity(D, _p(basis), len(full), _p(full.amp), _p(full.out), _p(full.inn),
                       _p(full.parity), row0, nrows, _p(counts), 0, 0, 0, 0, nthreads)
    rowptr = np.zeros(nrows + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(counts)
    nnz = int(rowptr[-1])
    cols = np.zeros(max(nnz, 1), dtype=np.int64)
    terms = np.zeros(max(nnz, 1), dtype=np.int32)
    negs = np.zeros(max(nnz, 1), dtype=np.int8)
    L.orc_connectivity(D, _p(basis), len(full), _p(full.amp), _p(full.out), _p(full.inn),
                       _p(full.parity), row0, nrows, 0, _p(rowptr), _p(cols), _p(terms),
                       _p(negs), nthreads)
    return rowptr, cols[:nnz], terms[:nnz], negs[:nnz]


def bracket2(bra_basis, bra, full2: TermList, full1: TermList, ket_basis, ket) -> complex:
    """multiplication.jl:170-183,214-223: <bra| op2 op1 |ket> over explicit term lists."""
    bb = np.ascontiguousarray(bra_basis, dtype=np.uint64)
    kb = np.ascontiguousarray(ket_basis, dtype=np.uint64)
    bv = np.ascontiguousarray(bra, dtype=np.complex128)
    kv = np.ascontiguousarray(ket, dtype=np.complex128)
    out = np.zeros(2)
    lib().orc_bracket2(len(bb), _p(bb), _p(bv), len(kb), _p(kb), _p(kv), len(full2),
                       _p(full2.amp), _p(full2.out), _p(full2.inn), _p(full2.parity), len(full1),
                       _p(full1.amp), _p(full1.out), _p(full1.inn), _p(full1.parity), _p(out))
    return complex(out[0], out[1])


def pes_block_rdm(basis, psi, A_states, ph_mask: int = 0) -> np.ndarray:
    """particle_reduced_density_matrix.jl:37-88: the particle reduced density matrix of psi
    on the part-A block A_states (nA x nA, Hermitian)."""
    b = np.ascontiguousarray(basis, dtype=np.uint64)
    v = np.ascontiguousarray(psi, dtype=np.complex128)
    A = np.ascontiguousarray(A_states, dtype=np.uint64)
    out = np.zeros((len(A), len(A)), dtype=np.complex128)
    lib().orc_pes_block_rdm(len(b), _p(b), _p(v), len(A), _p(A), int(ph_mask), _p(out))
    return out


def dense_matrix(basis: np.ndarray, full: TermList) -> np.ndarray:
    """Dense H from the oracle connectivity (small sectors only)."""
    D = len(basis)
    rowptr, cols, terms, negs = connectivity(basis, full)
    H = np.zeros((D, D), dtype=np.complex128)
    for i in range(D):
        for e in range(rowptr[i], rowptr[i + 1]):
            a = full.amp[terms[e]]
            H[i, cols[e]] += -a if negs[e] else a
    return H


# ------------------------------------------------------------- term-list restatement
def _scatter1(V: complex, i: int, j: int, upper: bool):
    """EDCore/src/types/scattering.jl:101-120 (1-based orbitals)."""
    im = 1 << (i - 1)
    jm = 1 << (j - 1)
    if upper:
        if jm < im:
            im, jm = jm, im
            V = V.conjugate()
        elif jm == im:
            V = complex(V.real, 0.0)
    return (V, im, jm)


def _scatter2(V: complex, i1: int, i2: int, j2: int, j1: int, upper: bool):
    """EDCore/src/types/scattering.jl:121-153 (1-based orbitals)."""
    im = (1 << (i1 - 1)) | (1 << (i2 - 1))
    jm = (1 << (j1 - 1)) | (1 << (j2 - 1))
    if j1 == j2 or i1 == i2:
        return (0j, im, jm)
    if i1 < i2:
        V = -V
    if j1 < j2:
        V = -V
    if upper:
        if jm < im:
            im, jm = jm, im
            V = V.conjugate()
        elif jm == im:
            V = complex(V.real, 0.0)
    return (V, im, jm)


def _sort_merge(terms):
    """EDCore/src/types/scattering.jl:258-282: stable sort by (in, out), sum duplicates."""
    terms = sorted(terms, key=lambda t: (t[2], t[1]))
    merged = []
    for V, o, i in terms:
        if merged and merged[-1][1] == o and merged[-1][2] == i:
            merged[-1] = (merged[-1][0] + V, o, i)
        else:
            merged.append((V, o, i))
    return merged


def _kcoords(k_list, Gk, shifts):
    """scat_list.jl:5-28 — k/G per dimension (exact Fractions for integer shifts)."""
    dim, nk = k_list.shape
    exact = all(isinstance(s, (int, Fraction)) for row in shifts for s in row)
    kc = [[None] * nk for _ in range(dim)]
    sh = [[None] * len(shifts[0]) for _ in range(dim)]
    for d in range(dim):
        g = Gk[d]
        for k in range(nk):
            v = Fraction(int(k_list[d, k])) if exact else float(k_list[d, k])
            kc[d][k] = v / g if g != 0 else v
        for c in range(len(shifts[0])):
            v = Fraction(shifts[d][c]) if exact else float(shifts[d][c])
            sh[d][c] = v / g if g != 0 else v
    return kc, sh


def scatterlist_onebody(k_list, Gk, Nc_mix, Nc_conserve, H_one: Callable, coordinate: bool,
                        kshift=None):
    """MomentumED/src/preparation/scat_list.jl:78-142. Returns [(amp, out, in)] sorted."""
    k_list = np.asarray(k_list)
    dim, Nk = k_list.shape
    Ncm, Ncc = Nc_mix, Nc_conserve
    shifts = [[0] * Ncc for _ in range(dim)] if kshift is None else kshift
    kc, sh = _kcoords(k_list, Gk, shifts)
    out = []
    for cc in range(1, Ncc + 1):
        for cmi in range(1, Ncm + 1):
            for cmf in range(1, Ncm + 1):
                cf = cmf + (cc - 1) * Ncm
                ci = cmi + (cc - 1) * Ncm
                for k in range(1, Nk + 1):
                    i_out = k + Nk * (cf - 1)
                    i_in = k + Nk * (ci - 1)
                    if i_in >= i_out:
                        if coordinate:
                            kk = tuple(kc[d][k - 1] + sh[d][cc - 1] for d in range(dim))
                            V = complex(H_one(kk, cf, ci))
                        else:
                            V = complex(H_one(k, cf, ci))
                        if V != 0:
                            out.append(_scatter1(V, i_out, i_in, True))
    return _sort_merge(out)


def scatterlist_twobody(k_list, Gk, Nc_mix, Nc_conserve, H_two: Callable, coordinate: bool,
                        kshift=None):
    """MomentumED/src/preparation/scat_list.jl:425-572, as a direct loop over momentum
    quadruples (ki1>=ki2, kf1>=kf2 with equal total momentum) instead of the reference's
    Dict of pair groups (:174-200); the per-tuple rules (:474-487) and amplitude (:496-509)
    are the reference's. Returns [(amp, out, in)] sorted."""
    k_list = np.asarray(k_list)
    dim, Nk = k_list.shape
    Ncm, Ncc = Nc_mix, Nc_conserve
    Nc = Ncm * Ncc
    sys_size = Nk if all(g != 0 for g in Gk) else 1
    shifts = [[0] * Ncc for _ in range(dim)] if kshift is None else kshift
    kc, sh = _kcoords(k_list, Gk, shifts)

    def ktot(a, b):
        return tuple((int(k_list[d, a - 1] + k_list[d, b - 1]) % Gk[d]) if Gk[d] != 0
                     else int(k_list[d, a - 1] + k_list[d, b - 1]) for d in range(dim))

    out = []
    for ki1 in range(1, Nk + 1):
        for ki2 in range(1, ki1 + 1):
            K = ktot(ki1, ki2)
            for kf1 in range(1, Nk + 1):
                for kf2 in range(1, kf1 + 1):
                    if ktot(kf1, kf2) != K:
                        continue
                    for ci1 in range(1, Nc + 1):
                        for ci2 in range(1, Nc + 1):
                            for cf1 in range(1, Nc + 1):
                                for cf2 in range(1, Nc + 1):
                                    f1 = kf1 + Nk * (cf1 - 1)
                                    f2 = kf2 + Nk * (cf2 - 1)
                                    i1 = ki1 + Nk * (ci1 - 1)
                                    i2 = ki2 + Nk * (ci2 - 1)
                                    if i1 == i2 or f1 == f2:
                                        continue
                                    if ki1 == ki2 and i1 < i2:
                                        continue
                                    if kf1 == kf2 and f1 < f2:
                                        continue
                                    if not ((min(i1, i2), max(i1, i2)) >= (min(f1, f2), max(f1, f2))):
                                        continue
                                    if coordinate:
                                        ccf1 = (cf1 - 1) // Ncm
                                        ccf2 = (cf2 - 1) // Ncm
                                        cci1 = (ci1 - 1) // Ncm
                                        cci2 = (ci2 - 1) // Ncm
                                        c_f1 = tuple(kc[d][kf1 - 1] + sh[d][ccf1] for d in range(dim))
                                        c_f2 = tuple(kc[d][kf2 - 1] + sh[d][ccf2] for d in range(dim))
                                        c_i1 = tuple(kc[d][ki1 - 1] + sh[d][cci1] for d in range(dim))
                                        c_i2 = tuple(kc[d][ki2 - 1] + sh[d][cci2] for d in range(dim))
                                        vd = complex(H_two(c_f1, c_f2, c_i2, c_i1, cf1, cf2, ci2, ci1))
                                        vx = complex(H_two(c_f1, c_f2, c_i1, c_i2, cf1, cf2, ci1, ci2))
                                    else:
                                        vd = complex(H_two(kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1))
                                        vx = complex(H_two(kf1, kf2, ki1, ki2, cf1, cf2, ci1, ci2))
                                    d = vd - vx
                                    amp = complex(d.real / sys_size, d.imag / sys_size)
                                    if amp != 0:
                                        out.append(_scatter2(amp, f1, f2, i2, i1, True))
    return _sort_merge(out)


def terms_to_list(terms) -> TermList:
    return TermList([t[0] for t in terms], [t[1] for t in terms], [t[2] for t in terms])


# ------------------------------------------------------------------- Jordan-Wigner oracle
def _jw_apply(ops, s: int):
    """Apply a product of fermion operators, rightmost first, to the occupation state s.
    ops = [(p, dagger)] with 0-based orbital p. Jordan-Wigner convention: c_p and c†_p carry
    the sign (-1)^(number of occupied orbitals below p). Returns (sign, state) or None."""
    sign = 1
    for p, dag in reversed(ops):
        occ = (s >> p) & 1
        if dag == occ:
            return None
        if bin(s & ((1 << p) - 1)).count("1") & 1:
            sign = -sign
        s ^= 1 << p
    return sign, s


def jw_hamiltonian(k_list, Gk, Nc_mix, Nc_conserve, H_one=None, H_two=None,
                   coordinate_one=True, coordinate_two=True):
    """Textbook second quantisation:
    H = sum H_one(k,cf,ci) c†_{k,cf} c_{k,ci} (cf, ci in the same conserved block)
      + 1/(2 N) sum_{momentum-conserving f1 f2 i1 i2} H_two(kf1,kf2,ki2,ki1,cf1,cf2,ci2,ci1)
                 c†_{f1} c†_{f2} c_{i2} c_{i1}
    with N = Nk when all Gk != 0 — the operator the reference's term lists encode for an
    interaction symmetric under vertex exchange and Hermitian. Returns a list of
    (coefficient, operator string) terms; use jw_sector_block to get a sector's matrix."""
    k_list = np.asarray(k_list)
    dim, Nk = k_list.shape
    Nc = Nc_mix * Nc_conserve
    sys_size = Nk if all(g != 0 for g in Gk) else 1
    kc, sh = _kcoords(k_list, Gk, [[0] * Nc_conserve for _ in range(dim)])

    def coord(k):
        return tuple(kc[d][k - 1] for d in range(dim))

    def orb(k, c):
        return k - 1 + Nk * (c - 1)
    terms = []
    if H_one is not None:
        for cc in range(1, Nc_conserve + 1):
            for cmf in range(1, Nc_mix + 1):
                for cmi in range(1, Nc_mix + 1):
                    cf = cmf + (cc - 1) * Nc_mix
                    ci = cmi + (cc - 1) * Nc_mix
                    for k in range(1, Nk + 1):
                        V = complex(H_one(coord(k) if coordinate_one else k, cf, ci))
                        if V != 0:
                            terms.append((V, [(orb(k, cf), 1), (orb(k, ci), 0)]))
    if H_two is not None:
        def ktot(a, b):
            return tuple((int(k_list[d, a - 1] + k_list[d, b - 1]) % Gk[d]) if Gk[d] != 0
                         else int(k_list[d, a - 1] + k_list[d, b - 1]) for d in range(dim))
        for kf1 in range(1, Nk + 1):
            for kf2 in range(1, Nk + 1):
                for ki1 in range(1, Nk + 1):
                    for ki2 in range(1, Nk + 1):
                        if ktot(kf1, kf2) != ktot(ki1, ki2):
                            continue
                        for cf1, cf2, ci1, ci2 in itertools.product(range(1, Nc + 1), repeat=4):
                            if coordinate_two:
                                V = H_two(coord(kf1), coord(kf2), coord(ki2), coord(ki1),
                                          cf1, cf2, ci2, ci1)
                            else:
                                V = H_two(kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1)
                            V = complex(V)
                            if V == 0:
                                continue
                            f1, f2 = orb(kf1, cf1), orb(kf2, cf2)
                            i1, i2 = orb(ki1, ci1), orb(ki2, ci2)
                            terms.append((V / (2 * sys_size),
                                          [(f1, 1), (f2, 1), (i2, 0), (i1, 0)]))
    return terms


def jw_sector_block(terms, basis: np.ndarray) -> np.ndarray:
    """Matrix of the Jordan-Wigner operator on the sector basis (columns = input states)."""
    basis = [int(v) for v in basis]
    index = {s: i for i, s in enumerate(basis)}
    H = np.zeros((len(basis), len(basis)), dtype=np.complex128)
    for col, s in enumerate(basis):
        for V, ops in terms:
            r = _jw_apply(ops, s)
            if r is None:
                continue
            row = index.get(r[1])
            if row is None:
                raise AssertionError("operator leaves the sector")
            H[row, col] += r[0] * V
    return H
