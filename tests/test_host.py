"""Host-side logic of the momentumed package (no GPU): the Scatter / MBOperator /
HilbertSubspace semantics against the reference's own unit tests, the term builders against
the oracle's restatement, argument validation, and the C-ABI library's exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import momentumed as me
from momentumed import (EDPara, ED_scatterlist_onebody, ED_scatterlist_twobody, HilbertSubspace,
                        MBOperator, MBS64, MBS64Vector, Scatter, get_body, index_fit,
                        isdiagonal, isupper, sort_merge_scatlist)
from momentumed import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def M(bits, *occ):
    return MBS64(bits, list(occ)).n


# ------------------------------------------------ EDCore/test/test_scattering.jl:1-69
def test_scatter_basic():
    s1 = Scatter(1.0, 1, 2, bits=10)
    assert s1.Amp == 1.0 and s1.out == M(10, 1) and s1.in_ == M(10, 2)
    assert get_body(s1) == (1, 1)
    s2 = Scatter(0.5, 1, 2, 4, 3, bits=10)
    assert s2.Amp == 0.5 and s2.out == M(10, 1, 2) and s2.in_ == M(10, 3, 4)
    assert get_body(s2) == (2, 2)
    assert s1.adjoint() == Scatter(1.0, 2, 1, bits=10)
    assert isdiagonal(Scatter(1.0, 1, 1, bits=10)) and not isdiagonal(s1)


def test_scatter_normal_ordering():
    s = Scatter(1.0 + 1.0j, 2, 1, bits=10, upper_hermitian=True)
    assert s.in_ == M(10, 2) and s.out == M(10, 1) and s.Amp == 1.0 - 1.0j
    s = Scatter(1.0 + 0.0j, 2, 1, 3, 4, bits=10, upper_hermitian=True)
    assert s.in_ == M(10, 4, 3) and s.out == M(10, 2, 1) and s.Amp == 1.0
    s = Scatter(1.0 + 0.0j, 1, 2, 3, 4, bits=10, upper_hermitian=True)
    assert s.in_ == M(10, 4, 3) and s.out == M(10, 2, 1) and s.Amp == -1.0


def test_scatter_operations_and_merge():
    s1 = Scatter(1.0, 1, 2, bits=10)
    s2 = Scatter(2.0, 1, 2, bits=10)
    s3 = Scatter(1.0, 2, 1, bits=10)
    s4 = Scatter(1.0, 1, 1, bits=10, upper_hermitian=True)
    assert s1 == s2 and s1 != s3 and s3 < s1
    assert (s1 + s2).Amp == 3.0 and (2 * s1).Amp == 2.0
    merged = sort_merge_scatlist([s1, s2], check_upper=False)
    assert len(merged) == 1 and merged[0].Amp == 3.0
    merged = sort_merge_scatlist([s1, s3], check_upper=False)
    assert len(merged) == 2 and merged[0] < merged[1]
    with pytest.raises(AssertionError):
        sort_merge_scatlist([s3], check_upper=True)
    assert sort_merge_scatlist([s4], check_upper=True) == [s4]


def test_scatter_nbody_normal_ordering():
    """N=3 path (scattering.jl:63-100): reordering creation/annihilation indices into
    descending order flips the sign once per transposition."""
    a = Scatter(1.0, 5, 3, 1, 2, 4, 6, bits=8)   # c†5 c†3 c†1 c6 c4 c2
    b = Scatter(1.0, 3, 5, 1, 2, 4, 6, bits=8)   # one swap among the creations
    c = Scatter(1.0, 3, 5, 1, 4, 2, 6, bits=8)   # and one among the annihilations
    assert a.Amp == 1.0 and b.Amp == -1.0 and c.Amp == 1.0
    assert a.out == M(8, 1, 3, 5) and a.in_ == M(8, 2, 4, 6)
    assert Scatter(1.0, 1, 1, 2, 3, 4, 5, bits=8).Amp == 0


# ------------------------------------------------ test_operator.jl, test_hilbert_subspace.jl
def test_mboperator_reference_vectors():
    s1 = Scatter(1.0 + 0.0j, 1, 2, bits=10)
    s2 = Scatter(0.5 + 0.0j, 1, 2, 4, 3, bits=10)
    op = MBOperator([s1], [s2], upper_hermitian=False)
    assert not isupper(op) and len(op.scats) == 2
    assert get_body(op.scats[0]) == (1, 1) and get_body(op.scats[1]) == (2, 2)
    with pytest.raises(AssertionError):
        MBOperator([Scatter(1.0, 2, 1, bits=4)], upper_hermitian=True)


def test_hilbert_subspace_reference_vectors():
    space = HilbertSubspace([i for i in range(5)], bits=8)
    assert len(space) == 5 and space.bits == 8
    assert space.get(2) == 3
    assert not index_fit(space.get(5), space, 5)
    v = np.arange(5) + 1j
    mv = MBS64Vector(v, space)
    assert mv.dot(mv) == pytest.approx(np.vdot(v, v))
    with pytest.warns(UserWarning):
        HilbertSubspace([3, 1, 2], bits=4)


def test_mbs64_reference_vectors():
    # EDCore/test/test_manybodystate_basis.jl:1-36
    assert MBS64(8, n=10).n == 10
    assert (MBS64(2, n=1) * MBS64(3, n=2)) == MBS64(5, n=10)
    assert me.occ_list(10) == [2, 4] and me.make_mask64([2, 4]) == 10
    with pytest.raises(AssertionError):
        MBS64(4, n=16)


def test_mbs64vector_operations():
    """manybodystate_vector.jl:63-117: length, size, similar (same subspace), dot with the
    first argument conjugated, + and - on one subspace only."""
    sp = HilbertSubspace([1, 2, 4], bits=3)
    a = me.MBS64Vector(np.array([1 + 1j, 2, 3j]), sp)
    b = me.MBS64Vector(np.array([1, 1j, 1]), sp)
    assert len(a) == 3 and a.size() == (3,)
    z = a.similar()
    assert z.space is sp and z.size() == (3,)
    assert a.dot(b) == np.vdot(a.vec, b.vec)
    np.testing.assert_array_equal((a + b).vec, a.vec + b.vec)
    np.testing.assert_array_equal((a - b).vec, a.vec - b.vec)
    other = me.MBS64Vector(np.zeros(3), HilbertSubspace([1, 2, 4], bits=3))
    with pytest.raises(AssertionError):
        a + other


def test_mbs64_manipulation_reference_vectors():
    """EDCore/test/test_manybodystate_basis.jl:1-46: bits, physicality, union, equality and
    order, occupation tests and updates, the sign count, flips."""
    mbs = MBS64(8, n=10)                                     # 00001010
    assert me.get_bits(mbs) == 8 and me.isphysical(mbs)
    assert not me.isphysical(me.core._raw(8, 256))
    assert MBS64(4, n=5) + MBS64(4, n=10) == MBS64(4, n=15)
    assert MBS64(8, n=10) == MBS64(8, n=10) and MBS64(8, n=10) != MBS64(8, n=11)
    assert MBS64(8, n=10) < MBS64(8, n=11)
    assert MBS64(8, [2, 4]) == mbs and MBS64(5, [1, 2], [1, 3, 5]) == MBS64(5, [1, 3])
    assert me.isoccupied(mbs, [2, 4]) and not me.isoccupied(mbs, [1, 2])
    assert me.isempty(mbs, [1, 3]) and not me.isempty(mbs, [2, 3])
    assert me.occupy(MBS64(8, n=0), [2, 4]) == mbs
    assert me.empty(mbs, [2, 4]) == MBS64(8, n=0)
    assert me.scat_occ_number(mbs, (2, 4)) == 0
    assert me.scat_occ_number(mbs, (1, 5)) == 2
    assert me.flip(mbs, [2, 4]) == MBS64(8, n=0)
    assert me.flip(mbs, [1, 5]) == MBS64(8, n=27)            # 00011011
    assert me.flip(MBS64(8, n=0), [2, 4]) == mbs
    with pytest.raises(AssertionError):
        me.occupy(mbs, [2, 3])
    with pytest.raises(AssertionError):
        me.empty(mbs, [1, 2])
    assert me.MBS64_complete(mbs) == MBS64(8, n=0b11110101)
    # mask forms take a raw bit mask and check nothing
    assert me.occupy(mbs, 1 << 8, check=True).n == 10 | 256
    assert me.isoccupied(mbs, 10) and me.isempty(mbs, 5)
    seen = []
    me.for_each_occ(seen.append, mbs)
    assert seen == [2, 4]


def test_scat_occ_number_matches_parity_mask(oracle_mod):
    """The sign count of a state and an operator set equals popcount(state & parity_mask)
    for every pair and single orbital of 10 bits (the single-orbital case: every bit above
    it), in the mask and the list forms."""
    rng = np.random.default_rng(5)
    for n in rng.integers(0, 1 << 10, 40):
        st = MBS64(10, n=int(n))
        for a in range(1, 11):
            want = bin(int(n) & oracle_mod.parity_mask(1 << (a - 1))).count("1")
            assert me.scat_occ_number(st, (a,)) == want == me.scat_occ_number(st, 1 << (a - 1))
            for b in range(a + 1, 11):
                m = (1 << (a - 1)) | (1 << (b - 1))
                want = bin(int(n) & oracle_mod.parity_mask(m)).count("1")
                assert me.scat_occ_number(st, [b, a]) == want == me.scat_occ_number(int(n), m)


def test_colex_iterators_reference_vectors():
    """EDCore/test/test_colex_mbslist.jl:4-79: colex order, edge cases t = 0, t = n,
    n = t = 0, sortedness; masked iteration; rank windows."""
    import math
    assert [m.n for m in me.ColexMBS64(4, 2)] == [MBS64(4, c).n for c in
                                                  [(1, 2), (1, 3), (2, 3), (1, 4), (2, 4),
                                                   (3, 4)]]
    assert list(me.ColexMBS64(3, 0)) == [MBS64(3, [])]
    assert list(me.ColexMBS64(3, 3)) == [MBS64(3, (1, 2, 3))]
    assert list(me.ColexMBS64(0, 0)) == [MBS64(0, [])]
    big = [m.n for m in me.ColexMBS64(5, 3)]
    assert len(big) == math.comb(5, 3) and big == sorted(big)
    assert list(me.ColexMBS64Mask(5, 2, [1, 3, 5])) == [MBS64(5, (1, 3)), MBS64(5, (1, 5)),
                                                         MBS64(5, (3, 5))]
    assert list(me.ColexMBS64Mask(5, 0, [1, 3, 5])) == [MBS64(5, [])]
    assert list(me.ColexMBS64Mask(5, 2, [2, 4])) == [MBS64(5, (2, 4))]
    lm = [m.n for m in me.ColexMBS64Mask(10, 3, [1, 2, 3, 4, 5, 6])]
    assert len(lm) == 20 and lm == sorted(lm)
    # every state of 12 bits with 5 set, in order, and windows of it from the unranking
    full = [m.n for m in me.ColexMBS64(12, 5)]
    assert full == [x for x in range(1 << 12) if bin(x).count("1") == 5]
    for lo, hi in [(1, 1), (7, 300), (790, 792), (792, 900)]:
        assert [m.n for m in me.ColexMBS64(12, 5, lo, hi)] == full[lo - 1:hi]
    assert list(me.ColexMBS64(12, 5, 800, 900)) == [] and len(me.ColexMBS64(4, 5)) == 0


def test_parity_mask_matches_oracle(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(500):
        v = int(rng.integers(0, 2 ** 63)) | (int(rng.integers(0, 2)) << 63)
        assert me.parity_mask(v) == oracle_mod.parity_mask(v)


# ------------------------------------------------ EDPara validation
def test_edpara_validation():
    k = np.zeros((2, 33), dtype=np.int64)
    with pytest.raises(AssertionError):
        EDPara(k_list=k, Gk=(1, 1), Nc_conserve=2)           # 66 orbitals > 64 bits
    with pytest.raises(AssertionError):
        EDPara(k_list=np.zeros((2, 3), dtype=np.int64), Gk=(-1, 1))
    p = EDPara(k_list=np.zeros((2, 3), dtype=np.int64), Gk=(1, 1))
    assert p.Nk == 3 and p.Nc == 1 and p.bits == 3
    assert me.momentum_residue((7, -1), (3, 0)) == (1, -1)


# ------------------------------------------------ term builders vs the oracle restatement
def _oracle_terms(O, para):
    t2 = O.scatterlist_twobody(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, para.H_two,
                               para.two_momentum_coordinate)
    t1 = []
    if para._H_one_given:
        t1 = O.scatterlist_onebody(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve,
                                   para.H_one, para.one_momentum_coordinate)
    return O._sort_merge(t1 + t2)


@pytest.mark.parametrize("name", ["small", "medium"])
def test_term_builder_matches_oracle(oracle_mod, name):
    from momentumed import workloads
    w = workloads.get(name)
    want = _oracle_terms(oracle_mod, w.para)
    got = w.operator().scats
    assert len(got) == len(want)
    for s, (a, o, i) in zip(got, want):
        assert (s.out, s.in_) == (o, i) and s.Amp == a


def test_term_builder_multicomponent_matches_oracle(oracle_mod):
    from momentumed.workloads import CoulombGaussian, mesh
    k = mesh(2, 3)

    def h1(kc, cf, ci):
        return 0.25 + 0.5j if (cf, ci) == (1, 2) else (0.25 - 0.5j if (cf, ci) == (2, 1) else 0j)
    para = EDPara(k_list=k, Gk=(2, 3), Nc_mix=2, Nc_conserve=2, H_one=h1,
                  H_two=CoulombGaussian(k, (2, 3), perturb=0.02, seed=5),
                  two_momentum_coordinate=False)
    want = _oracle_terms(oracle_mod, para)
    got = MBOperator(ED_scatterlist_onebody(para), ED_scatterlist_twobody(para),
                     upper_hermitian=True).scats
    assert [(s.Amp, s.out, s.in_) for s in got] == [tuple(t) for t in want]


def test_kshift_coordinates_are_exact():
    from fractions import Fraction
    seen = []
    para = EDPara(k_list=np.array([[0, 1, 2], [0, 0, 0]]), Gk=(3, 1),
                  H_one=lambda k, cf, ci: seen.append(k) or 1.0)
    ED_scatterlist_onebody(para, kshift=(1, 0))
    assert seen[1] == (Fraction(2, 3), Fraction(0))


# ------------------------------------------------ solve argument handling (no GPU needed)
def test_edsolve_argument_errors():
    sp = HilbertSubspace([1, 2], bits=2)
    op = MBOperator([Scatter(1.0, 1, 2, bits=2, upper_hermitian=True)], upper_hermitian=True)
    with pytest.raises(AssertionError):
        me.EDsolve(sp, op, N=0)
    with pytest.raises(NotImplementedError):
        me.EDsolve(sp, op, N=1, method="dense", device="cpu")
    with pytest.raises(NotImplementedError):
        me.EDsolve(sp, op, N=1, device="cpu", method_info=False)
    with pytest.raises(NotImplementedError):
        me.EDsolve(sp, op, N=1, method="sparse", device="cpu")
    non_upper = MBOperator([Scatter(1.0, 1, 2, bits=2)], upper_hermitian=False)
    with pytest.raises(AssertionError):
        me.EDsolve(sp, non_upper, N=1, method_info=False)
    with pytest.raises(ValueError):
        me.EDsolve(sp, [Scatter(1.0, 1, 1, bits=2)], N=1, method="map")
    # element_type as the reference accepts it (sparse_matrix.jl:72): Float64 or Float32 only,
    # refused before any device work
    for bad in (int, np.complex128, "Float16"):
        with pytest.raises(AssertionError, match="element_type"):
            me.EDsolve(sp, op, N=1, element_type=bad, method_info=False)


# ------------------------------------------------ the C-ABI library
def test_library_exports_every_declared_symbol():
    L = _lib.load()
    header = open(os.path.join(ROOT, "include", "edm.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(edm_[a-z0-9_]+)\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert isinstance(getattr(L, name), ctypes._CFuncPtr)


def test_library_reports_errors_without_gpu():
    L = _lib.load()
    # bad system description: no device work is needed to reject it
    import momentumed.device as dev
    sysd = dev.system_struct(np.zeros((2, 40), dtype=np.int64), (1, 1), 1, 2, [1, 1])
    n = ctypes.c_int64()
    rc = L.edm_num_sectors(ctypes.byref(sysd), ctypes.byref(n))
    assert rc == _lib.EDM_EINVAL
    assert b"64 bits" in L.edm_last_error(None)
    sysd = dev.system_struct(np.array([[0, 1, 2], [0, 0, 0]]), (3, 0), 1, 1, [2])
    assert L.edm_num_sectors(ctypes.byref(sysd), ctypes.byref(n)) == 0
    assert n.value == 3 * 1          # Gk=0 direction: range [0*2, 0*2] -> one value


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "momentumed.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".jl")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "liboracle" not in src and "orc_" not in src, f


def _eigh_lib(A):
    import ctypes
    from momentumed import _lib
    L = _lib.load()
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    w = np.empty(n)
    Z = np.empty((n, n))
    _lib.check(L.edm_projected_eigh(n, A.ctypes.data, w.ctypes.data, Z.ctypes.data))
    return w, Z            # Z[i] = eigenvector i


@pytest.mark.parametrize("kind", ["random", "arrow", "tridiag", "degenerate", "diag", "tiny"])
def test_projected_eigh_matches_lapack(kind):
    """The host Rayleigh-Ritz eigensolver against LAPACK (numpy.linalg.eigh): eigenvalues to
    1e-13 of ||A||, orthonormal vectors, residuals A z = w z to 1e-13 of ||A||."""
    rng = np.random.default_rng({"random": 1, "arrow": 2, "tridiag": 3, "degenerate": 4,
                                 "diag": 5, "tiny": 6}[kind])
    mats = []
    for n in (1, 2, 3, 7, 30, 40):
        if kind == "random":
            B = rng.normal(size=(n, n)); A = B + B.T
        elif kind == "arrow":          # kept Ritz values + coupling row + tridiagonal tail
            A = np.zeros((n, n)); k = (3 * n) // 5
            A[np.arange(n), np.arange(n)] = rng.normal(size=n)
            for c in range(min(k, n - 1)):
                A[c, k] = A[k, c] = rng.normal() * 1e-3
            for i in range(k, n - 1):
                A[i, i + 1] = A[i + 1, i] = abs(rng.normal())
        elif kind == "tridiag":
            A = np.diag(rng.normal(size=n)) + np.diag(np.abs(rng.normal(size=n - 1)), 1)
            A = A + np.triu(A, 1).T
        elif kind == "degenerate":     # repeated eigenvalues
            Qm, _ = np.linalg.qr(rng.normal(size=(n, n)))
            ev = np.repeat(rng.normal(size=(n + 2) // 3), 3)[:n]
            A = (Qm * ev) @ Qm.T; A = 0.5 * (A + A.T)
        elif kind == "diag":
            A = np.diag(rng.permutation(n).astype(float))
        else:
            A = np.diag(rng.normal(size=n) * 1e-200)
        mats.append(A)
    for A in mats:
        w, Z = _eigh_lib(A)
        ref = np.linalg.eigvalsh(A)
        scale = max(np.abs(A).max(), 1e-300)
        assert np.all(np.diff(w) >= 0)
        np.testing.assert_allclose(w, ref, rtol=0, atol=1e-13 * scale * A.shape[0])
        np.testing.assert_allclose(Z @ Z.T, np.eye(A.shape[0]), atol=1e-13)
        assert np.abs(A @ Z.T - Z.T * w).max() <= 1e-13 * scale * A.shape[0]


def _heigh_lib(A):
    from momentumed import _lib
    L = _lib.load()
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.complex128)
    w = np.empty(n)
    Z = np.empty((n, n), dtype=np.complex128)
    _lib.check(L.edm_projected_heigh(n, A.ctypes.data, w.ctypes.data, Z.ctypes.data))
    return w, Z            # Z[i] = eigenvector i


@pytest.mark.parametrize("kind", ["random", "block_arrow", "tridiag", "degenerate", "real",
                                  "diag", "tiny"])
def test_projected_heigh_matches_lapack(kind):
    """The complex Hermitian Rayleigh-Ritz eigensolver (block Lanczos) against LAPACK
    (numpy.linalg.eigh, zheevd): eigenvalues to 1e-13 of ||A||, orthonormal vectors, residuals
    A z = w z to 1e-13 of ||A||. A real input must give the real solver's eigenvalues."""
    rng = np.random.default_rng({"random": 11, "block_arrow": 12, "tridiag": 13,
                                 "degenerate": 14, "real": 15, "diag": 16, "tiny": 17}[kind])

    def crand(*shape):
        return rng.normal(size=shape) + 1j * rng.normal(size=shape)
    mats = []
    for n in (1, 2, 3, 8, 32, 44):
        if kind == "random":
            B = crand(n, n); A = B + B.conj().T
        elif kind == "block_arrow":    # kept Ritz values, a coupling block, block tridiagonal
            A = np.zeros((n, n), dtype=complex); b = 4; k = max(0, n - 3 * b)
            A[np.arange(k), np.arange(k)] = rng.normal(size=k)
            for j in range(k, n, b):
                e = min(j + b, n)
                C = crand(e - j, e - j); A[j:e, j:e] = C + C.conj().T
                if e < n:
                    R = np.triu(crand(min(b, n - e), e - j))
                    A[e:e + R.shape[0], j:e] = R; A[j:e, e:e + R.shape[0]] = R.conj().T
            if k and k < n:
                S = crand(min(b, n - k), k) * 1e-3
                A[k:k + S.shape[0], :k] = S; A[:k, k:k + S.shape[0]] = S.conj().T
        elif kind == "tridiag":
            A = np.diag(rng.normal(size=n)).astype(complex)
            off = crand(n - 1)
            A += np.diag(off, -1) + np.diag(off.conj(), 1)
        elif kind == "degenerate":     # repeated eigenvalues
            Qm, _ = np.linalg.qr(crand(n, n))
            ev = np.repeat(rng.normal(size=(n + 3) // 4), 4)[:n]
            A = (Qm * ev) @ Qm.conj().T; A = 0.5 * (A + A.conj().T)
        elif kind == "real":
            B = rng.normal(size=(n, n)); A = (B + B.T).astype(complex)
        elif kind == "diag":
            A = np.diag(rng.permutation(n).astype(float)).astype(complex)
        else:
            A = (np.diag(rng.normal(size=n)) + 1e-3 * np.eye(n, k=1) + 1e-3 * np.eye(n, k=-1)
                 ).astype(complex) * 1e-200
        mats.append(A)
    for A in mats:
        n = A.shape[0]
        w, Z = _heigh_lib(A)
        ref = np.linalg.eigvalsh(A)
        scale = max(np.abs(A).max(), 1e-300)
        assert np.all(np.diff(w) >= 0)
        np.testing.assert_allclose(w, ref, rtol=0, atol=1e-13 * scale * n)
        np.testing.assert_allclose(Z.conj() @ Z.T, np.eye(n), atol=1e-13)
        assert np.abs(A @ Z.T - Z.T * w).max() <= 1e-13 * scale * n
        if kind == "real":
            wr, _ = _eigh_lib(A.real)
            np.testing.assert_allclose(w, wr, rtol=0, atol=1e-13 * scale * n)


# ------------------------------------------------ batched / served workloads (host side)
def test_served_stream_rules():
    from fractions import Fraction
    from momentumed import serve as S
    reqs = S.served_stream(rate=50.0, duration=20.0, seed=7)
    assert reqs == S.served_stream(rate=50.0, duration=20.0, seed=7)    # seeded
    assert 800 < len(reqs) < 1200                                       # Poisson(1000)
    assert all(a.arrival < b.arrival for a, b in zip(reqs, reqs[1:]))
    seen = set()
    for r in reqs:
        nk = r.mesh[0] * r.mesh[1]
        assert r.mesh in S.MESHES and r.filling in (Fraction(1, 3), Fraction(2, 5))
        assert r.N == nk * r.filling                 # redraw rule: N is always an integer
        assert 0 <= r.sector[0] < r.mesh[0] and 0 <= r.sector[1] < r.mesh[1]
        assert 1.0 <= r.screening < 4.0
        seen.add(r.label)
    # 4x5@1/3, 4x6@2/5 and 6x6@2/5 never appear; the other seven pairs do
    assert seen == {"3x5@1/3", "3x5@2/5", "4x5@2/5", "4x6@1/3", "5x6@1/3", "5x6@2/5", "6x6@1/3"}


def test_batched_workload_dimensions(oracle_mod):
    """The batched sweep's sectors span the intended 1.2k-20k range (oracle enumeration)."""
    from momentumed import workloads as W
    r = W.batched_ratios()
    assert len(r) == 64 and r.min() >= 0.5 and r.max() <= 2.0
    np.testing.assert_array_equal(r, W.batched_ratios())
    para = W.batched_para(1.0)
    dims = []
    for f in W.BATCH_FILLINGS:
        lists, ks = oracle_mod.enumerate_sectors(para.k_list, para.Gk, 1, 2, list(f))
        assert len(lists) == 12
        dims += [len(l) for l in lists]
    assert min(dims) == 1210 and max(dims) == 20439


def test_problem_bytes_accounts_for_the_krylov_workspace():
    from momentumed.batch import problem_bytes
    for D, nev, kd in ((1000, 4, 30), (20439, 4, 30), (5, 2, 30), (34770492, 6, 40)):
        m = min(max(kd, nev + 2), D)
        b = problem_bytes(D, nev, kd)
        assert b >= (2 * m + 1) * D * 16
        assert b <= (2 * m + 1) * (D + 64) * 16 + 4 * (m + 1) * ((D + 1023) // 1024 + 2) * 16 + 65536
    assert problem_bytes(2000, 4, 30) > problem_bytes(1000, 4, 30)


def test_gauge_fixing_closed_forms():
    from momentumed import ED_connection_gaugefixing
    sp = HilbertSubspace([3, 5, 6], bits=3)
    v = me.MBS64Vector(np.array([1j, 2.0, -1.0]), sp)
    ED_connection_gaugefixing(v)
    np.testing.assert_allclose(v.vec, [1.0, -2j, 1j], atol=1e-15)
    a = me.MBS64Vector(np.array([1j, 0.0, 0.0]), sp)
    b = me.MBS64Vector(np.array([0.0, 1.0, 0.0]), sp)
    ED_connection_gaugefixing([a, b])                # det [[1j, 0], [0, 1]] = 1j
    np.testing.assert_allclose(a.vec, [1.0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(b.vec, [0, 1.0, 0], atol=1e-15)


def _brute_sectors(para, N_each, mask=None, rule=None):
    """All states of the whole bit space, filtered component by component: |N_c| set bits
    inside component c (inside the mask, when given), momentum residue of the components'
    residues with holes negated, then the rule; grouped by momentum, ascending."""
    from momentumed.preparation import _momentum_list, momentum_residue
    width = para.Nk * para.Nc_mix
    out = {}
    for s in range(1 << para.bits):
        ok, K = True, [0] * para.dim
        for c, n in enumerate(N_each):
            field = (s >> (c * width)) & ((1 << width) - 1)
            if bin(field).count("1") != abs(n):
                ok = False
                break
            if mask is not None and any(((field >> b) & 1) and (c * width + b + 1) not in mask
                                        for b in range(width)):
                ok = False
                break
            kc = [0] * para.dim
            for b in range(width):
                if field >> b & 1:
                    for d in range(para.dim):
                        kc[d] += int(para.k_list[d, b % para.Nk])
            kc = momentum_residue(kc, para.Gk)
            sgn = (n > 0) - (n < 0)
            K = [K[d] + sgn * kc[d] for d in range(para.dim)]
        if not ok or (rule is not None and not rule(MBS64(para.bits, n=s))):
            continue
        out.setdefault(momentum_residue(K, para.Gk), []).append(s)
    order = _momentum_list(para, N_each)
    return [(k, out[k]) for k in order if k in out]


def test_momentum_subspaces_mask_holes_selection():
    """ED_momentum_subspaces with mask, holes (negative N) and selection_rule
    (momentum_decomposition.jl:101-152,188-262) against brute force over every state: one
    component with a mask, two holes, a selection rule; two conserved components with a
    particle and a hole component under a mask that spans both."""
    from momentumed.workloads import mesh
    p1 = EDPara(k_list=mesh(3, 4), Gk=(3, 4))
    p2 = EDPara(k_list=mesh(2, 3), Gk=(2, 3), Nc_conserve=2)
    cases = [(p1, [3], [1, 2, 4, 7, 8, 11, 12], None),
             (p1, [-2], None, None),
             (p1, [3], None, lambda m: not (m.n & 0b101)),
             (p2, [2, -1], [1, 2, 3, 5, 8, 9, 10, 12], None),
             (p2, [1, 2], None, lambda m: bin(m.n).count("1") == 3 and m.n % 3 != 0)]
    for para, N, mask, rule in cases:
        subs, ks = me.ED_momentum_subspaces(para, tuple(N), mask=mask, selection_rule=rule)
        want = _brute_sectors(para, N, mask, rule)
        assert ks == [k for k, _ in want], (N, mask)
        for sp, (_, states) in zip(subs, want):
            assert [int(x) for x in sp.list] == states
        assert sum(len(s) for _, s in want) > 0


def test_host_enumeration_equals_the_oracle(oracle_mod):
    """The host enumeration (forced by an always-true selection rule) reproduces the
    oracle's pinned sector lists bit for bit: two conserved components, and a mixed one."""
    from momentumed.workloads import mesh
    for para, N in ((EDPara(k_list=mesh(2, 3), Gk=(2, 3), Nc_conserve=2), [2, 3]),
                    (EDPara(k_list=mesh(1, 5), Gk=(1, 5), Nc_mix=2), [4])):
        subs, ks = me.ED_momentum_subspaces(para, tuple(N), selection_rule=lambda m: True)
        lists, oks = oracle_mod.enumerate_sectors(para.k_list, para.Gk, para.Nc_mix,
                                                  para.Nc_conserve, N)
        assert ks == oks
        for sp, ob in zip(subs, lists):
            np.testing.assert_array_equal(sp.list, ob)


@pytest.mark.parametrize("kind", ["random", "krylov_schur", "triangular", "hermitian", "jordan",
                                  "diag", "tiny"])
def test_projected_schur_matches_lapack(kind):
    """The complex Schur decomposition of the general eigensolver (edm_projected_schur)
    against numpy (LAPACK zgeev): A = Z T Z^H to 1e-13 of ||A||, Z unitary, T upper
    triangular with its diagonal ascending by real part (then imaginary part), the diagonal
    equal to LAPACK's eigenvalues. 'krylov_schur' is the projected matrix after a restart
    (upper triangular block, a full coupling row, Hessenberg below); 'jordan' a defective
    matrix (one eigenvalue, one eigenvector), where the eigenvalues agree to ~eps^(1/n)."""
    from momentumed import _lib
    L = _lib.load()
    rng = np.random.default_rng({"random": 21, "krylov_schur": 22, "triangular": 23,
                                 "hermitian": 24, "jordan": 25, "diag": 26, "tiny": 27}[kind])

    def crand(*shape):
        return rng.normal(size=shape) + 1j * rng.normal(size=shape)
    for n in (1, 2, 3, 5, 12, 30, 64):
        if kind == "random":
            A = crand(n, n)
        elif kind == "krylov_schur":
            k = (3 * n) // 5
            A = np.triu(crand(n, n), -1)
            A[:k, :k] = np.triu(crand(k, k))
            if k < n:
                A[k, :k] = crand(k)
        elif kind == "triangular":
            A = np.triu(crand(n, n))
        elif kind == "hermitian":
            B = crand(n, n)
            A = B + B.conj().T
        elif kind == "jordan":
            A = np.eye(n) * (0.3 + 0.7j) + np.diag(np.ones(n - 1), 1)
        elif kind == "diag":
            A = np.diag(crand(n))
        else:
            A = crand(n, n) * 1e-300
        A = np.ascontiguousarray(A, dtype=np.complex128)
        T = np.zeros_like(A)
        Z = np.zeros_like(A)
        assert L.edm_projected_schur(n, A.ctypes.data, T.ctypes.data, Z.ctypes.data) == 0
        scale = max(np.abs(A).max(), 1e-300)
        assert np.abs(Z @ T @ Z.conj().T - A).max() <= 1e-13 * scale
        assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-13
        assert np.all(np.tril(T, -1) == 0)
        d = np.diag(T)
        for i in range(n - 1):
            assert d[i].real < d[i + 1].real or (d[i].real == d[i + 1].real
                                                 and d[i].imag <= d[i + 1].imag)
        if kind == "jordan":
            assert np.abs(d - (0.3 + 0.7j)).max() <= 1e-15 ** (1.0 / n) * 10
            continue
        ev = np.linalg.eigvals(A)
        ev = ev[np.lexsort((ev.imag, ev.real))]
        assert np.abs(d - ev).max() <= 1e-10 * scale
