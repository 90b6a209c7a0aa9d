"""Preparation: system parameters, momentum sectors and term lists.

Mirror of MomentumED/src/preparation/ (init_parameter.jl, momentum_decomposition.jl,
scat_list.jl). Momentum and component indices passed to the H_one / H_two callbacks are
1-based, exactly as the reference passes them. The basis of a momentum sector is enumerated on
the device (libedm.so); term lists are built on the host, where the user's callbacks live.
"""
from __future__ import annotations

import itertools
from fractions import Fraction
from typing import Callable, Sequence

import numpy as np

from .core import ColexMBS64, ColexMBS64Mask, HilbertSubspace, MBS64, Scatter, sort_merge_scatlist

__all__ = [
    "EDPara", "momentum_residue", "MBS_totalmomentum", "ED_momentum_subspaces",
    "ED_scatterlist_onebody", "ED_scatterlist_twobody", "ED_sortedScatterList_onebody",
    "ED_sortedScatterList_twobody", "group_momentum_pairs",
]


def _zero(*_args, **_kw):
    return 0j


class EDPara:
    """init_parameter.jl:62-146.

    EDPara(k_list=..., Gk=..., Nc_mix=1, Nc_conserve=1, H_one=None, H_two=None,
           one_momentum_coordinate=None, two_momentum_coordinate=None)

    k_list is a dim x Nk integer matrix (column k = momentum of index k+1). Julia detects the
    callback signature from its method table; Python cannot, so a callback taking momentum
    indices instead of coordinates must say so with *_momentum_coordinate=False (the default,
    None, means coordinates, which is what the reference picks for untyped callbacks)."""

    def __init__(self, *, k_list, Gk: Sequence[int], Nc_mix: int = 1, Nc_conserve: int = 1,
                 H_one: Callable | None = None, H_two: Callable | None = None,
                 one_momentum_coordinate: bool | None = None,
                 two_momentum_coordinate: bool | None = None):
        k_list = np.asarray(k_list, dtype=np.int64)
        if k_list.ndim == 1:
            k_list = k_list.reshape(1, -1)
        self.k_list = k_list
        self.Gk = tuple(int(g) for g in Gk)
        if len(self.Gk) != k_list.shape[0]:
            raise AssertionError("Gk must have one entry per momentum dimension")
        self.dim = k_list.shape[0]
        self.Nk = k_list.shape[1]
        self.Nc_mix = int(Nc_mix)
        self.Nc_conserve = int(Nc_conserve)
        self.Nc = self.Nc_mix * self.Nc_conserve
        if self.Nc <= 0:
            raise AssertionError("Number of components must be positive")
        if self.Nk * self.Nc > 64:
            raise AssertionError("The total orbital number exceeds 64 bits.")
        if any(g < 0 for g in self.Gk):
            raise AssertionError("Gk must all be non-negative integers.")
        self.H_one = H_one if H_one is not None else _zero
        self.H_two = H_two if H_two is not None else _zero
        self.one_momentum_coordinate = True if one_momentum_coordinate is None else bool(one_momentum_coordinate)
        self.two_momentum_coordinate = True if two_momentum_coordinate is None else bool(two_momentum_coordinate)
        self._H_one_given = H_one is not None
        self._H_two_given = H_two is not None

    @property
    def bits(self) -> int:
        return self.Nk * self.Nc


def momentum_residue(k: Sequence[int], Gk: Sequence[int]) -> tuple:
    """init_parameter.jl:148-151."""
    return tuple(int(k[d]) if Gk[d] == 0 else int(k[d]) % Gk[d] for d in range(len(Gk)))


def MBS_totalmomentum(para: EDPara, mbs) -> tuple:
    """momentum_decomposition.jl:49-77: total momentum of a state (MBS64 or int), or of a
    tuple of 1-based orbital indices (repeats counted)."""
    k = [0] * para.dim
    if isinstance(mbs, tuple):
        idx = list(mbs)
    else:
        n = mbs.n if isinstance(mbs, MBS64) else int(mbs)
        idx = []
        while n:
            low = n & -n
            idx.append(low.bit_length())
            n ^= low
    for i in idx:
        ik = (i - 1) % para.Nk
        for d in range(para.dim):
            k[d] += int(para.k_list[d, ik])
    return momentum_residue(k, para.Gk)


def _momentum_list(para: EDPara, N_each) -> list[tuple]:
    """momentum_decomposition.jl:209-219: product of per-dim ranges, first coordinate fastest."""
    ntot = int(sum(N_each))
    ranges = []
    for d in range(para.dim):
        if para.Gk[d] == 0:
            ranges.append(range(int(para.k_list[d].min()) * ntot, int(para.k_list[d].max()) * ntot + 1))
        else:
            ranges.append(range(para.Gk[d]))
    # first coordinate fastest (Iterators.product order)
    return [tuple(reversed(t)) for t in itertools.product(*reversed(ranges))]


def _wanted_momenta(para: EDPara, N_each, momentum_restriction: bool, momentum_list):
    if momentum_restriction:
        wanted = []
        for k in (momentum_list or []):
            k = momentum_residue(k, para.Gk)
            if k not in wanted:
                wanted.append(k)
        return wanted
    return _momentum_list(para, N_each)


def _host_momentum_subspaces(para: EDPara, N_each, momentum_restriction: bool, momentum_list,
                             mask, selection_rule):
    """momentum_decomposition.jl:101-152,188-262 with a mask, holes or a selection rule (the
    forms the reference's entanglement analysis uses), enumerated on the host: component c
    (the last one in the highest bits) holds |N_c| particles, or |N_c| holes when N_c < 0,
    on its orbitals (those of `mask` only, when given: 1-based orbital indices of the whole
    state); its momentum is the residue of its orbitals' momenta, negated for holes; a state
    is kept when selection_rule(state) holds, in the sector of its total momentum residue.
    Every list comes out ascending (outer components vary slowest, each one in colex order)."""
    width = para.Nk * para.Nc_mix
    if mask is not None:
        mask = sorted(int(o) for o in mask)
    wanted = _wanted_momenta(para, N_each, momentum_restriction, momentum_list)
    pats, moms = [], []
    for c, n in enumerate(N_each):
        if not abs(n) <= width:
            raise AssertionError("Invalid number of electrons in one component")
        if mask is None:
            it = ColexMBS64(width, abs(n))
        else:
            local = [o - c * width for o in mask if c * width < o <= (c + 1) * width]
            it = ColexMBS64Mask(width, abs(n), local)
        p = np.array([st.n for st in it], dtype=np.uint64)
        sgn = (n > 0) - (n < 0)
        m = np.array([[sgn * v for v in MBS_totalmomentum(para, int(x))] for x in p],
                     dtype=np.int64).reshape(len(p), para.dim)
        pats.append(p)
        moms.append(m)
    # outer products, the last component outermost: ascending states
    states, ksum = pats[0], moms[0]
    for c in range(1, len(N_each)):
        hi = pats[c] << np.uint64(c * width)
        states = (hi[:, None] | states[None, :]).reshape(-1)
        ksum = (moms[c][:, None, :] + ksum[None, :, :]).reshape(-1, para.dim)
    res = np.stack([ksum[:, d] % para.Gk[d] if para.Gk[d] else ksum[:, d]
                    for d in range(para.dim)], axis=1) if len(states) else ksum
    keep = np.ones(len(states), dtype=bool)
    if selection_rule is not None:
        keep = np.array([bool(selection_rule(MBS64(para.bits, n=int(x)))) for x in states],
                        dtype=bool)
    subspaces, ks = [], []
    for k in wanted:
        sel = keep & np.all(res == np.asarray(k, dtype=np.int64), axis=1)
        if sel.any():
            subspaces.append(HilbertSubspace(states[sel], bits=para.bits))
            ks.append(k)
    return subspaces, ks


def ED_momentum_subspaces(para: EDPara, N_each_component, *, momentum_restriction: bool = False,
                          momentum_list: list | None = None, device: str = "rocm",
                          mask=None, selection_rule=None):
    """momentum_decomposition.jl:188-262: the Hilbert subspace of every momentum sector, in
    momentum_list order with empty sectors dropped. Returns (subspaces, subspace_k).

    The states are enumerated on the device; each returned HilbertSubspace keeps its device
    copy and fetches its host list lazily. The forms with `mask`, `selection_rule` or holes
    (negative N), which the reference uses to build entanglement-analysis blocks, are
    enumerated on the host (as the reference does for every form) and uploaded when a device
    operation first needs them."""
    from . import device as dev
    if isinstance(N_each_component, (int, np.integer)):
        N_each_component = (int(N_each_component),)
    N_each = [int(n) for n in N_each_component]
    if len(N_each) != para.Nc_conserve:
        raise AssertionError("The length of number_list must be equal to the number of "
                             f"conserved components {para.Nc_conserve}")
    if mask is not None or selection_rule is not None or any(n < 0 for n in N_each):
        return _host_momentum_subspaces(para, N_each, momentum_restriction, momentum_list,
                                        mask, selection_rule)
    if device not in ("rocm", "gpu"):
        raise ValueError("the basis is enumerated on the device (device='rocm')")
    dev.throw_gpu_unavailable("rocm")
    ctx = dev.get_context(0)
    system = dev.system_struct(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, N_each)
    mlist = _momentum_list(para, N_each)
    if momentum_restriction:
        wanted = _wanted_momenta(para, N_each, True, momentum_list)
        sectors = [mlist.index(k) for k in wanted if k in mlist]
    else:
        sectors = list(range(len(mlist)))
    subspaces, ks = [], []
    if momentum_restriction:
        bases = [dev.DeviceBasis.enumerate(ctx, system, s) for s in sectors]
    else:   # every sector from one pass over the occupation space
        bases = dev.DeviceBasis.enumerate_all(ctx, system)
        if len(bases) != len(mlist):
            raise AssertionError("device sector count differs from momentum_list")
    for s, b in zip(sectors, bases):
        if b.D == 0:
            b.close()
            continue
        sp = HilbertSubspace(None, bits=para.bits, _device=b, _length=b.D)
        sp.sector_info = (system, s)
        subspaces.append(sp)
        ks.append(mlist[s])
    return subspaces, ks


# ------------------------------------------------------------------------- term lists
def _shift_decipher(kshift, Ncc: int, dim: int):
    """scat_list.jl:29-50 -> shifts[d][cc] (0-based cc)."""
    if kshift is None:
        return [[0] * Ncc for _ in range(dim)]
    if isinstance(kshift, tuple):
        return [[kshift[d]] * Ncc for d in range(dim)]
    if isinstance(kshift, list):
        if len(kshift) != Ncc:
            raise AssertionError("length of kshift isn't equal to para.Nc_conserve.")
        return [[kshift[c][d] for c in range(Ncc)] for d in range(dim)]
    arr = np.asarray(kshift)
    if arr.shape != (dim, Ncc):
        raise AssertionError("size of kshift isn't equal to para.Nc_conserve.")
    return [[arr[d, c].item() for c in range(Ncc)] for d in range(dim)]


def _coordinates(para: EDPara, shifts):
    """scat_list.jl:5-28: coordinates k/Gk — exact Fractions for integer/rational shifts,
    floats otherwise."""
    exact = all(isinstance(v, (int, np.integer, Fraction)) for row in shifts for v in row)
    kc = [[None] * para.Nk for _ in range(para.dim)]
    sh = [[None] * len(shifts[0]) for _ in range(para.dim)]
    for d in range(para.dim):
        g = para.Gk[d]
        for k in range(para.Nk):
            v = Fraction(int(para.k_list[d, k])) if exact else float(para.k_list[d, k])
            kc[d][k] = v / g if g != 0 else v
        for c in range(len(shifts[0])):
            v = Fraction(shifts[d][c]) if exact else float(shifts[d][c])
            sh[d][c] = v / g if g != 0 else v
    return kc, sh


def _div(z: complex, n: int) -> complex:
    return complex(z.real / n, z.imag / n)


def ED_scatterlist_onebody(para: EDPara, *, kshift=None, **H_one_kw) -> list[Scatter]:
    """scat_list.jl:78-142: one-body terms of H_one(k, cf, ci), upper-Hermitian, sorted."""
    Nk, Ncc, Ncm, bits = para.Nk, para.Nc_conserve, para.Nc_mix, para.bits
    terms = []
    if para.one_momentum_coordinate:
        kc, sh = _coordinates(para, _shift_decipher(kshift, Ncc, para.dim))
    for cc in range(1, Ncc + 1):
        for cmi in range(1, Ncm + 1):
            for cmf in range(1, Ncm + 1):
                cf = cmf + (cc - 1) * Ncm
                ci = cmi + (cc - 1) * Ncm
                for k in range(1, Nk + 1):
                    i_out = k + Nk * (cf - 1)
                    i_in = k + Nk * (ci - 1)
                    if i_in < i_out:
                        continue
                    if para.one_momentum_coordinate:
                        kk = tuple(kc[d][k - 1] + sh[d][cc - 1] for d in range(para.dim))
                        V = complex(para.H_one(kk, cf, ci, **H_one_kw))
                    else:
                        V = complex(para.H_one(k, cf, ci, **H_one_kw))
                    if V != 0:
                        terms.append(Scatter(V, i_out, i_in, bits=bits, upper_hermitian=True))
    return sort_merge_scatlist(terms)


def group_momentum_pairs(para: EDPara) -> dict:
    """scat_list.jl:174-200: momentum-index pairs (i >= j) grouped by total momentum, in
    first-seen order."""
    groups: dict = {}
    for i in range(1, para.Nk + 1):
        for j in range(1, i + 1):
            K = MBS_totalmomentum(para, (i, j))
            groups.setdefault(K, []).append((i, j))
    return groups


def ED_scatterlist_twobody(para: EDPara, *, kshift=None, **H_two_kw) -> list[Scatter]:
    """scat_list.jl:425-572: antisymmetrised two-body terms
    amp = (H_two(kf1,kf2,ki2,ki1,...) - H_two(kf1,kf2,ki1,ki2,...)) / sys_size over momentum-
    conserving pairs, with the reference's skip rules (:474-487) and upper-Hermitian
    canonicalisation; zero amplitudes are dropped; the result is sorted and merged."""
    Nk, Nc, Ncm, Ncc, bits = para.Nk, para.Nc, para.Nc_mix, para.Nc_conserve, para.bits
    sys_size = Nk if all(g != 0 for g in para.Gk) else 1
    coord = para.two_momentum_coordinate
    if coord:
        kc, sh = _coordinates(para, _shift_decipher(kshift, Ncc, para.dim))

        def kcoord(k, c):
            cc = (c - 1) // Ncm
            return tuple(kc[d][k - 1] + sh[d][cc] for d in range(para.dim))
    H2 = para.H_two
    terms = []
    for _K, pairs in group_momentum_pairs(para).items():
        for (ki1, ki2) in pairs:
            for (kf1, kf2) in pairs:
                for ci1 in range(1, Nc + 1):
                    for ci2 in range(1, Nc + 1):
                        i1 = ki1 + Nk * (ci1 - 1)
                        i2 = ki2 + Nk * (ci2 - 1)
                        if i1 == i2 or (ki1 == ki2 and i1 < i2):
                            continue
                        lo_i, hi_i = (i1, i2) if i1 < i2 else (i2, i1)
                        for cf1 in range(1, Nc + 1):
                            for cf2 in range(1, Nc + 1):
                                f1 = kf1 + Nk * (cf1 - 1)
                                f2 = kf2 + Nk * (cf2 - 1)
                                if f1 == f2 or (kf1 == kf2 and f1 < f2):
                                    continue
                                lo_f, hi_f = (f1, f2) if f1 < f2 else (f2, f1)
                                if (lo_i, hi_i) < (lo_f, hi_f):
                                    continue
                                if coord:
                                    a_f1, a_f2 = kcoord(kf1, cf1), kcoord(kf2, cf2)
                                    a_i1, a_i2 = kcoord(ki1, ci1), kcoord(ki2, ci2)
                                    vd = complex(H2(a_f1, a_f2, a_i2, a_i1, cf1, cf2, ci2, ci1, **H_two_kw))
                                    vx = complex(H2(a_f1, a_f2, a_i1, a_i2, cf1, cf2, ci1, ci2, **H_two_kw))
                                else:
                                    vd = complex(H2(kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1, **H_two_kw))
                                    vx = complex(H2(kf1, kf2, ki1, ki2, cf1, cf2, ci1, ci2, **H_two_kw))
                                amp = _div(vd - vx, sys_size)
                                if amp != 0:
                                    terms.append(Scatter(amp, f1, f2, i2, i1, bits=bits,
                                                         upper_hermitian=True))
    return sort_merge_scatlist(terms)


# The older names of the two term builders (aliases left commented out in
# scat_list.jl:574-575), exported for scripts that still use them.
ED_sortedScatterList_onebody = ED_scatterlist_onebody
ED_sortedScatterList_twobody = ED_scatterlist_twobody
