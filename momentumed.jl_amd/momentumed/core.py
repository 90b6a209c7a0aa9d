"""Core algebra: occupation bitstrings, scattering terms, operators, subspaces, vectors.

Host-side mirror of EDCore (EDCore/src/EDCore.jl:3-11). Orbital indices are 1-based as in
the reference (orbital i is bit i-1 of a uint64 state). The Julia functions that end in "!"
(occupy!, empty!, flip!) return new states in the reference too and are named here without
the "!". Brackets and other eigenvector consumers (EDCore/src/multiplication.jl:153-307) run
on the device and live in analysis.py.
"""
from __future__ import annotations

from typing import Iterable, Sequence

import numpy as np

__all__ = [
    "MBS64", "make_mask64", "occ_list", "parity_mask", "Scatter", "isupper", "isdiagonal",
    "get_body", "sort_merge_scatlist", "MBOperator", "HilbertSubspace", "MBS64Vector",
    "index_fit", "get_bits", "isphysical", "MBS64_complete", "isoccupied", "isempty",
    "occupy", "empty", "flip", "scat_occ_number", "ColexMBS64", "ColexMBS64Mask",
    "for_each_occ",
]


# ---------------------------------------------------------------------------- states
def make_mask64(occ: Iterable[int]) -> int:
    """EDCore/src/types/manybodystate_basis.jl:161-182: mask with bits i-1 set, 1 <= i <= 64."""
    m = 0
    for i in occ:
        if not 1 <= i <= 64:
            raise AssertionError("orbital index must be in 1..64")
        m |= 1 << (i - 1)
    return m


def occ_list(n: int) -> list[int]:
    """manybodystate_basis.jl:131-140: 1-based occupied orbitals, ascending."""
    out = []
    n = int(n)
    while n:
        low = n & -n
        out.append(low.bit_length())
        n ^= low
    return out


def for_each_occ(f, mbs) -> None:
    """manybodystate_basis.jl:142-149: f(i) for every occupied 1-based orbital, ascending."""
    for i in occ_list(mbs.n if isinstance(mbs, MBS64) else mbs):
        f(i)


class MBS64:
    """manybodystate_basis.jl:24-32: a 64-bit occupation state with `bits` physical bits."""

    __slots__ = ("n", "bits")

    def __init__(self, bits: int, occ=None, mask=None, *, n: int | None = None):
        """MBS64(bits, occ) sets the 1-based orbitals in occ; MBS64(bits, occ, mask) the
        orbitals mask[o - 1] for o in occ (manybodystate_basis.jl:195-209)."""
        if not 0 <= bits <= 64:
            raise AssertionError("The number of bits must be an integer between 0 and 64.")
        if n is None:
            occ = list(occ or [])
            if mask is not None:
                occ = [mask[o - 1] for o in occ]
            occ = sorted(occ)
            if len(set(occ)) != len(occ):
                raise AssertionError("Duplicate orbital indices when constructing MBS64.")
            n = make_mask64(occ)
        if n > (1 << bits) - 1:
            raise AssertionError("State representation out of range for given bits")
        self.n = int(n)
        self.bits = bits

    def __eq__(self, other):
        if isinstance(other, MBS64):
            return self.bits == other.bits and self.n == other.n
        return self.n == other

    def __lt__(self, other):
        return self.n < other.n

    def __hash__(self):
        return hash(self.n)

    def __mul__(self, other: "MBS64") -> "MBS64":
        """Concatenation (manybodystate_basis.jl:81-83): self takes the higher bits."""
        return MBS64(self.bits + other.bits, n=(self.n << other.bits) | other.n)

    def __add__(self, other: "MBS64") -> "MBS64":
        """Union of the occupations of two states of one space (manybodystate_basis.jl:
        92-94); a doubly occupied orbital counts once."""
        if self.bits != other.bits:
            raise TypeError("+ needs two MBS64 of the same bits")
        return MBS64(self.bits, n=self.n | other.n)

    def __repr__(self):
        s = format(self.n, f"0{max(self.bits, 1)}b")[::-1][: self.bits]
        return f"MBS64: {self.n} = {s} ({self.bits} bits)"


def _mask_of(i) -> int:
    """An orbital selection as a bit mask: an int is a mask already, a list or tuple holds
    1-based orbitals."""
    if isinstance(i, (int, np.integer)):
        return int(i)
    return make_mask64(i)


def get_bits(x) -> int:
    """manybodystate_basis.jl:53-55 (and the subspace / vector forms): physical bits."""
    if isinstance(x, MBS64Vector):
        return x.space.bits
    return x.bits


def isphysical(mbs: MBS64) -> bool:
    """manybodystate_basis.jl:62-69: no bit at or above `bits` is set."""
    return mbs.bits == 64 or mbs.n < (1 << mbs.bits)


def MBS64_complete(mbs: MBS64) -> MBS64:
    """manybodystate_basis.jl:216-218: every physical orbital flipped."""
    return MBS64(mbs.bits, n=((1 << mbs.bits) - 1) ^ mbs.n)


def isoccupied(mbs: MBS64, i) -> bool:
    """manybodystate_basis.jl:228-238: every orbital of i (an int mask, or a list or tuple
    of 1-based orbitals) occupied."""
    m = _mask_of(i)
    return mbs.n & m == m


def isempty(mbs: MBS64, i) -> bool:
    """manybodystate_basis.jl:256-266: every orbital of i empty."""
    return mbs.n & _mask_of(i) == 0


def occupy(mbs: MBS64, i, check: bool = True) -> MBS64:
    """occupy! (manybodystate_basis.jl:285-301): a new state with the orbitals of i
    occupied. The mask form checks nothing; the list form checks that they were empty
    (unless check=False) and that the result fits in `bits`."""
    if isinstance(i, (int, np.integer)):
        return _raw(mbs.bits, mbs.n | int(i))
    m = make_mask64(i)
    if check and mbs.n & m:
        raise AssertionError("Some orbitals are already occupied.")
    return MBS64(mbs.bits, n=mbs.n | m)


def empty(mbs: MBS64, i, check: bool = True) -> MBS64:
    """empty! (manybodystate_basis.jl:326-342): a new state with the orbitals of i emptied;
    the list form checks that they were occupied (unless check=False)."""
    if isinstance(i, (int, np.integer)):
        return _raw(mbs.bits, mbs.n & ~int(i) & ((1 << 64) - 1))
    m = make_mask64(i)
    if check and mbs.n & m != m:
        raise AssertionError("Some orbitals are already empty.")
    return MBS64(mbs.bits, n=mbs.n & ~m)


def flip(mbs: MBS64, i) -> MBS64:
    """flip! (manybodystate_basis.jl:365-375): occupations of the orbitals of i flipped."""
    if isinstance(i, (int, np.integer)):
        return _raw(mbs.bits, mbs.n ^ int(i))
    return MBS64(mbs.bits, n=mbs.n ^ make_mask64(i))


def _raw(bits: int, n: int) -> MBS64:
    """A state without the range check (the reference's reinterpret of the mask forms)."""
    out = MBS64.__new__(MBS64)
    out.n, out.bits = int(n), bits
    return out


def scat_occ_number(mbs, i) -> int:
    """manybodystate_basis.jl:397-459: the occupied orbitals of mbs (a state or a raw int)
    that a product of operators on the orbitals of i passes. The set bits of i pair up from
    the lowest; a pair covers the bits strictly between its two; an unpaired top bit covers
    every bit above it."""
    n = mbs.n if isinstance(mbs, MBS64) else int(mbs)
    if not isinstance(i, (int, np.integer)):
        if any(not 1 <= o <= (mbs.bits if isinstance(mbs, MBS64) else 64) for o in i):
            raise AssertionError("Invalid bit positions")
    bits = occ_list(_mask_of(i))
    cover = 0
    for p in range(0, len(bits), 2):
        lo = bits[p]                                        # 1-based: bit lo - 1
        hi = bits[p + 1] - 1 if p + 1 < len(bits) else 64   # exclusive upper bit
        cover |= ((1 << hi) - 1) & ~((1 << lo) - 1)
    return bin(n & cover).count("1")


class ColexMBS64:
    """colex_mbslist.jl:2-95: the states of n bits with t set, in colex order (= ascending
    as integers), optionally only the ranks start..end (1-based, inclusive). Iterates by the
    next-same-popcount bit step (Gosper's rule) from the first state, unranked by the
    combinatorial number system."""

    def __init__(self, n: int, t: int, start: int = 1, end: int | None = None):
        import math
        if not 0 <= n <= 64:
            raise AssertionError("0 <= n <= 64")
        if start < 1:
            raise AssertionError("start_number > 0")
        full = math.comb(n, t) if 0 <= t <= n else 0
        end = full if end is None else min(end, full)
        self.n, self.t, self.start, self.end = n, t, start, end

    def __len__(self):
        return max(0, self.end - self.start + 1)

    def _unrank(self, k: int) -> int:
        """Bit pattern of 0-based colex rank k."""
        import math
        x, state = k, 0
        for i in range(self.t, 0, -1):
            c = i - 1
            while math.comb(c + 1, i) <= x:
                c += 1
            state |= 1 << c
            x -= math.comb(c, i)
        return state

    def _patterns(self):
        if len(self) == 0:
            return
        x = self._unrank(self.start - 1)
        for _ in range(len(self)):
            yield x
            if x == 0:
                return
            low = x & -x
            ripple = x + low
            x = (((ripple ^ x) >> 2) // low) | ripple

    def __iter__(self):
        for x in self._patterns():
            yield MBS64(self.n, n=x)


class ColexMBS64Mask(ColexMBS64):
    """colex_mbslist.jl:96-172: the same over the orbitals of `mask` (1-based, sorted,
    unique): combinations of t of them in colex order of their positions in the mask, as
    states of n bits (ascending, since the mask is)."""

    def __init__(self, n: int, t: int, mask, start: int = 1, end: int | None = None):
        self.mask = [int(o) for o in mask]
        if any(b <= a for a, b in zip(self.mask, self.mask[1:])):
            raise AssertionError("the mask must be sorted and unique")
        if any(not 1 <= o <= n for o in self.mask):
            raise AssertionError("mask orbitals must lie in 1..n")
        super().__init__(len(self.mask), t, start, end)
        self.bits_total = n

    def __iter__(self):
        for x in self._patterns():
            yield MBS64(self.bits_total, [p + 1 for p in range(self.n) if x >> p & 1],
                        self.mask)


# ---------------------------------------------------------------------------- terms
def parity_mask(occ: int) -> int:
    """EDCore/src/types/scattering.jl:7-18 (an unpaired top bit p covers bits p+1..63)."""
    mask = 0
    occ = int(occ)
    M = (1 << 64) - 1
    while occ:
        tz1 = (occ & -occ).bit_length() - 1
        occ &= occ - 1
        if occ:
            tz2 = (occ & -occ).bit_length() - 1
            occ &= occ - 1
            mask += 1 << tz2
        mask -= 1 << (tz1 + 1) if tz1 + 1 < 64 else 0
    return mask & M


def _perm_parity(seq: Sequence[int]) -> int:
    inv = 0
    for a in range(len(seq)):
        for b in range(a + 1, len(seq)):
            if seq[a] > seq[b]:
                inv += 1
    return inv & 1


class Scatter:
    """A scattering term Amp * c†_i1 ... c†_iN c_jN ... c_j1 with masks `out` (creation) and
    `in` (annihilation) — EDCore/src/types/scattering.jl:52-62.

    Scatter(V, i, j, bits=..., upper_hermitian=False)             one-body (:101-120)
    Scatter(V, i1, i2, j2, j1, bits=..., upper_hermitian=False)   two-body (:121-153)
    Scatter(V, *out_in, bits=..., upper_hermitian=False)          N-body (:63-100)
    Orbitals are 1-based; the operator order is the reference's (creation indices first,
    annihilation indices in reverse).
    """

    __slots__ = ("Amp", "out", "in_", "parity_mask", "bits")

    def __init__(self, V, *out_in: int, bits: int, upper_hermitian: bool = False,
                 _raw: tuple | None = None):
        self.bits = bits
        if _raw is not None:
            self.Amp, self.out, self.in_ = complex(_raw[0]), int(_raw[1]), int(_raw[2])
            self.parity_mask = parity_mask(self.out) ^ parity_mask(self.in_)
            return
        V = complex(V)
        if len(out_in) % 2:
            raise AssertionError("number conservation requires annihilated and created "
                                 "indices number being even")
        N = len(out_in) // 2
        if N == 1:
            i, j = out_in
            im, jm = make_mask64((i,)), make_mask64((j,))
        elif N == 2:
            i1, i2, j2, j1 = out_in
            im, jm = make_mask64((i1, i2)), make_mask64((j1, j2))
            if j1 == j2 or i1 == i2:
                V = 0j
                upper_hermitian = False
            else:
                if i1 < i2:
                    V = -V
                if j1 < j2:
                    V = -V
        else:
            is_ = list(out_in[:N])
            js = list(out_in[2 * N - 1:N - 1:-1])
            im, jm = make_mask64(is_), make_mask64(js)
            if bin(im).count("1") < N or bin(jm).count("1") < N:
                V = 0j
                upper_hermitian = False
            else:
                i_sort = sorted(range(N), key=lambda k: -is_[k])
                j_sort = sorted(range(N), key=lambda k: -js[k])
                if (_perm_parity(i_sort) + _perm_parity(j_sort)) & 1:
                    V = -V
        if upper_hermitian:
            if jm < im:
                im, jm = jm, im
                V = V.conjugate()
            elif jm == im:
                V = complex(V.real, 0.0)
        if (im | jm) >> bits:
            raise AssertionError("State representation out of range for given bits")
        self.Amp = V
        self.out = im
        self.in_ = jm
        self.parity_mask = parity_mask(im) ^ parity_mask(jm)

    # `in` is a Python keyword; expose it under its reference name too
    def __getattr__(self, name):
        if name == "in":
            return self.in_
        raise AttributeError(name)

    def adjoint(self) -> "Scatter":
        """scattering.jl:185-187."""
        return Scatter(None, bits=self.bits, _raw=(self.Amp.conjugate(), self.in_, self.out))

    def key(self):
        """Ordering key of scattering.jl:205-207: (in, out)."""
        return (self.in_, self.out)

    def __eq__(self, other):
        return isinstance(other, Scatter) and self.in_ == other.in_ and self.out == other.out

    def __lt__(self, other):
        return self.key() < other.key()

    def __add__(self, other):
        if self != other:
            raise AssertionError("Can only add identical Scatter terms")
        return Scatter(None, bits=self.bits, _raw=(self.Amp + other.Amp, self.out, self.in_))

    def __rmul__(self, a):
        return Scatter(None, bits=self.bits, _raw=(complex(a) * self.Amp, self.out, self.in_))

    def __repr__(self):
        return (f"Scatter: c†_out {occ_list(self.out)[::-1]} c_in {occ_list(self.in_)}, "
                f"Amp = {self.Amp}")


def isupper(x) -> bool:
    """scattering.jl:170 for a Scatter; operator.jl:46 for an MBOperator."""
    if isinstance(x, MBOperator):
        return x.upper_hermitian
    return x.in_ > x.out or (x.in_ == x.out and x.Amp.imag == 0)


def isdiagonal(s: Scatter) -> bool:
    return s.in_ == s.out


def get_body(s: Scatter):
    return (bin(s.out).count("1"), bin(s.in_).count("1"))


def sort_merge_scatlist(*lists, check_upper: bool = True) -> list[Scatter]:
    """scattering.jl:258-282: stable sort by (in, out) and sum the amplitudes of equal terms."""
    terms = [s for lst in lists for s in lst]
    if not terms:
        return []
    if check_upper and not all(isupper(s) for s in terms):
        raise AssertionError("All Scatter terms must be in the upper triangular.")
    terms = sorted(terms, key=Scatter.key)
    merged = [0.0 * terms[0]]
    for s in terms:
        if s == merged[-1]:
            merged[-1] = merged[-1] + s
        else:
            merged.append(s)
    return merged


class MBOperator:
    """EDCore/src/types/operator.jl:24-32: a sorted, merged term list; `upper_hermitian` means
    the list holds one triangle of a Hermitian operator."""

    def __init__(self, *scats, upper_hermitian: bool):
        if len(scats) == 1 and scats[0] and isinstance(scats[0][0], (list, tuple)):
            scats = tuple(scats[0])
        self.scats = sort_merge_scatlist(*scats, check_upper=upper_hermitian)
        self.upper_hermitian = upper_hermitian

    def arrays(self):
        """Structure-of-arrays view: amp (complex128), out, in (uint64) — the layout
        create_gpu_linearmap flattens to (MomentumEDCUDAExt.jl:97-100)."""
        amp = np.array([s.Amp for s in self.scats], dtype=np.complex128)
        out = np.array([s.out for s in self.scats], dtype=np.uint64)
        inn = np.array([s.in_ for s in self.scats], dtype=np.uint64)
        return amp, out, inn

    @property
    def bits(self) -> int:
        return self.scats[0].bits if self.scats else 0

    def __mul__(self, other):
        """op * vec (EDCore/src/multiplication.jl:123-129): the operator applied to an
        MBS64Vector, as a new vector on the same subspace (targets outside it are dropped; an
        upper-Hermitian list acts with its adjoint half), computed by the device matvec.
        op2 * op1 is the product that applies op1 and then op2 to a vector (:131-140)."""
        if isinstance(other, MBOperator):
            return _OperatorProduct(self, other)
        if isinstance(other, MBS64Vector):
            if not self.scats:
                return MBS64Vector(np.zeros(len(other.vec), dtype=np.complex128), other.space)
            from .device import create_gpu_linearmap
            H = create_gpu_linearmap(self, other.space)
            return MBS64Vector(H(np.ascontiguousarray(other.vec, dtype=np.complex128)),
                               other.space)
        return NotImplemented

    def __repr__(self):
        h = "(Hermitian, upper triangle entries only) " if self.upper_hermitian else ""
        return f"Many Body Operator {h}:\n\t{len(self.scats)} scattering terms"


class _OperatorProduct:
    """op2 * op1 waiting for its vector: (op2 * op1) * vec applies op1, then op2, each on the
    vector's subspace (multiplication.jl:131-140)."""

    def __init__(self, op2: MBOperator, op1: MBOperator):
        self.op2, self.op1 = op2, op1

    def __mul__(self, vec):
        if not isinstance(vec, MBS64Vector):
            return NotImplemented
        return self.op2 * (self.op1 * vec)


# ---------------------------------------------------------------------------- spaces
class HilbertSubspace:
    """EDCore/src/types/hilbert_subspace.jl:19-30: a sorted basis of uint64 states.

    `list` is a numpy uint64 array. A subspace produced on the device keeps a handle to its
    device copy (`_device`) and fetches the host list only when it is read."""

    def __init__(self, states=None, *, bits: int, _device=None, _length: int | None = None):
        self.bits = bits
        self._device = _device
        self._list = None
        if states is not None:
            arr = np.ascontiguousarray(np.asarray(states, dtype=np.uint64))
            if arr.size > 1 and not np.all(arr[1:] > arr[:-1]):
                import warnings
                warnings.warn("The input list for HilbertSubspace is not sorted. "
                              "It will be sorted now.")
                arr = np.unique(arr)
            self._list = arr
            self.length = int(arr.size)
        else:
            self.length = int(_length)
        # (system, sector) when the subspace is a complete momentum sector
        self.sector_info = None

    @property
    def list(self) -> np.ndarray:
        if self._list is None:
            self._list = self._device.download()
        return self._list

    def __len__(self):
        return self.length

    def get(self, mbs) -> int:
        """hilbert_subspace.jl:67-69: searchsortedfirst (1-based)."""
        n = mbs.n if isinstance(mbs, MBS64) else int(mbs)
        return int(np.searchsorted(self.list, np.uint64(n), side="left")) + 1

    def __repr__(self):
        return f"Hilbert subspace {{bits = {self.bits}, dim = {self.length}}}"


def index_fit(index: int, space: HilbertSubspace, mbs) -> bool:
    """hilbert_subspace.jl:70-72."""
    n = mbs.n if isinstance(mbs, MBS64) else int(mbs)
    return 0 < index <= space.length and int(space.list[index - 1]) == n


class MBS64Vector:
    """EDCore/src/types/manybodystate_vector.jl:21-29: coefficients bound to a subspace."""

    def __init__(self, vec, space: HilbertSubspace):
        vec = np.asarray(vec)
        if vec.dtype != np.complex64:   # ComplexF32 coefficients stay single precision
            vec = vec.astype(np.complex128, copy=False)
        if vec.shape != (space.length,):
            raise AssertionError("vector length mismatches Hilbert space dimension.")
        self.vec = vec
        self.space = space

    def __len__(self):
        return self.vec.size

    def size(self) -> tuple:
        """manybodystate_vector.jl:70."""
        return (self.vec.size,)

    def similar(self) -> "MBS64Vector":
        """manybodystate_vector.jl:77-79: a vector on the same subspace (zeros here; the
        reference leaves the values undefined)."""
        return MBS64Vector(np.zeros_like(self.vec), self.space)

    def dot(self, other: "MBS64Vector") -> complex:
        """manybodystate_vector.jl:100-103 (conjugates the first argument)."""
        if self.space is not other.space:
            raise AssertionError("mbs vectors are not in the same subspace.")
        return complex(np.vdot(self.vec, other.vec))

    def __add__(self, other):
        if self.space is not other.space:
            raise AssertionError("mbs vectors are not in the same subspace.")
        return MBS64Vector(self.vec + other.vec, self.space)

    def __sub__(self, other):
        if self.space is not other.space:
            raise AssertionError("mbs vectors are not in the same subspace.")
        return MBS64Vector(self.vec - other.vec, self.space)
