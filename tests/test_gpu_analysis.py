"""Eigenvector consumers on the device against the oracle and closed forms.

Bars: brackets within 1e-12 of max scale of the oracle's dense matrices (between two
different sectors too); <psi|H|psi> of an eigenvector = its energy within 1e-10; the one-body
density matrix of a basis state = its occupations exactly, Hermitian with trace N for any
normalised state, entry by entry equal to the oracle's <psi|c_i^+ c_j|psi> within 1e-13;
step inner products equal vdot for unit orbital overlaps and the phase of the summed angles
for unimodular ones.
"""
import numpy as np
import pytest

import momentumed as me
from momentumed import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_full(O, op):
    amp, out, inn = op.arrays()
    return O.full_list(O.TermList(amp, out, inn), op.upper_hermitian)


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v / np.linalg.norm(v)


def test_reference_unit_vectors_on_device(gpu_ctx):
    """The reference's own unit tests, through the device path: EDCore/test/
    test_multiplication.jl:39-61 (Scatter(1+1im, 1, 2; upper_hermitian) on |01> gives
    (1-1im)|10>, on |10> gives (1+1im)|01>) with the device matvec, and :63-78
    (<10| c_2^+ c_1 |01> = 1) with the device bracket, one- and two-operator forms."""
    space = me.HilbertSubspace([0b01, 0b10], bits=2)
    s1 = me.Scatter(1.0 + 1.0j, 1, 2, bits=2, upper_hermitian=True)
    H = me.create_gpu_linearmap(me.LinearMap(me.MBOperator([s1], upper_hermitian=True),
                                             space), "rocm")
    y = H(np.array([1, 0], dtype=np.complex128))
    assert y[0] == 0 and y[1] == 1 - 1j
    y = H(np.array([0, 1], dtype=np.complex128))
    assert y[0] == 1 + 1j and y[1] == 0
    op = me.MBOperator([me.Scatter(1.0, 2, 1, bits=2)], upper_hermitian=False)
    bra = me.MBS64Vector(np.array([0, 1], dtype=np.complex128), space)   # <10|
    ket = me.MBS64Vector(np.array([1, 0], dtype=np.complex128), space)   # |01>
    assert abs(me.ED_bracket(bra, op, ket) - 1.0) < 1e-15
    ident = me.MBOperator([me.Scatter(1.0, 1, 1, bits=2), me.Scatter(1.0, 2, 2, bits=2)],
                          upper_hermitian=False)          # the particle number, 1 here
    assert abs(me.ED_bracket(bra, ident, op, ket) - 1.0) < 1e-15


def test_operator_times_vector(gpu_ctx, oracle_mod):
    """op * vec and op2 * op1 * vec (multiplication.jl:123-140) on the device: an
    upper-Hermitian operator, a non-Hermitian momentum-conserving one, and the product of
    the two, each against the oracle's dense matrices on the vector's subspace within 1e-13
    of the scale."""
    w = W.small()
    H = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][2]
    rng = np.random.default_rng(31)
    # the operator's first 200 terms with random complex amplitudes and no adjoints
    G = me.MBOperator([me.Scatter(None, bits=15,
                                  _raw=(complex(rng.normal(), rng.normal()), t.out, t.in_))
                       for t in H.scats[:200]], upper_hermitian=False)
    x = me.MBS64Vector(_rand(sp.length, 32), sp)
    Hd = oracle_mod.dense_matrix(sp.list, _oracle_full(oracle_mod, H))
    Gd = oracle_mod.dense_matrix(sp.list, _oracle_full(oracle_mod, G))
    assert np.abs(Gd - Gd.conj().T).max() > 1e-3          # not Hermitian
    for got, want in ((H * x, Hd @ x.vec), (G * x, Gd @ x.vec), (G * H * x, Gd @ (Hd @ x.vec))):
        assert got.space is sp
        assert np.abs(got.vec - want).max() <= 1e-13 * np.abs(want).max()


def test_reference_integration_example(gpu_ctx, oracle_mod):
    """The reference's integration test (MomentumED/test/example1_FQH_Laughlin.jl:93-111):
    ED_momentum_subspaces, ED_scatterlist_twobody and EDsolve(subspace, scat; N=10) in every
    one of the 12 sectors of the 3x4 mesh with 4 electrons (the scatter-list form switches to
    the dense method at these sizes, as in the reference). The test asserts nothing; here every
    sector's 10 energies equal the oracle's dense spectrum within 1e-10 and the vectors are
    eigenvectors (residual < 1e-9)."""
    import test_oracle as T
    para = T._example1_para()
    subs, ks = me.ED_momentum_subspaces(para, (4,))
    scat = me.ED_scatterlist_twobody(para)
    full = _oracle_full(oracle_mod, me.MBOperator(scat, upper_hermitian=True))
    assert len(subs) == 12
    for sp, k in zip(subs, ks):
        E, V = me.EDsolve(sp, scat, N=10, method_info=False)
        Hd = oracle_mod.dense_matrix(sp.list, full)
        ref = np.linalg.eigvalsh(Hd)[:10]
        np.testing.assert_allclose(E, ref, rtol=1e-10, atol=1e-12, err_msg=str(k))
        for e, v in zip(E, V):
            assert np.linalg.norm(Hd @ v.vec - e * v.vec) < 1e-9
    print(f"[example1] 12 sectors, lowest {min(float(np.min(me.EDsolve(sp, scat, N=1, method_info=False)[0])) for sp in subs):.10f}",
          flush=True)


def test_bracket_same_space(gpu_ctx, oracle_mod):
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][1]
    Hd = oracle_mod.dense_matrix(sp.list, _oracle_full(oracle_mod, op))
    a, b = _rand(sp.length, 1), _rand(sp.length, 2)
    got = me.ED_bracket(me.MBS64Vector(a, sp), op, me.MBS64Vector(b, sp))
    ref = np.vdot(a, Hd @ b)
    assert abs(got - ref) <= 1e-12 * np.abs(Hd).max() * sp.length
    E, V = me.EDsolve(sp, op, N=3, method="map", method_info=False)
    for e, v in zip(E, V):
        assert abs(me.ED_bracket(v, op, v) - e) <= 1e-10 * max(1.0, abs(e))


def test_bracket_between_sectors(gpu_ctx, oracle_mod):
    """A one-body operator that moves momentum, between two different sectors: against the
    oracle's dense matrix on the union of the two bases."""
    w = W.small()
    subs, ks = me.ED_momentum_subspaces(w.para, 5)
    # c_a^+ c_b with k_a - k_b = (1, 0): maps sector K to K + (1, 0)
    terms = []
    for b in range(1, 16):
        kb = ((b - 1) % 3, (b - 1) // 3)
        ka = ((kb[0] + 1) % 3, kb[1])
        a = ka[0] + 3 * ka[1] + 1
        terms.append(me.Scatter(complex(0.3 * b, -0.1 * b), a, b, bits=15))
    op = me.MBOperator(terms, upper_hermitian=False)
    ket_sp = subs[ks.index((0, 0))]
    bra_sp = subs[ks.index((1, 0))]
    union = np.union1d(bra_sp.list, ket_sp.list).astype(np.uint64)
    Hd = oracle_mod.dense_matrix(union, _oracle_full(oracle_mod, op))
    ri = np.searchsorted(union, bra_sp.list)
    ci = np.searchsorted(union, ket_sp.list)
    block = Hd[np.ix_(ri, ci)]
    assert np.abs(block).max() > 0
    a, b = _rand(bra_sp.length, 3), _rand(ket_sp.length, 4)
    got = me.ED_bracket(me.MBS64Vector(a, bra_sp), op, me.MBS64Vector(b, ket_sp))
    ref = np.vdot(a, block @ b)
    assert abs(got - ref) <= 1e-12 * np.abs(block).max() * union.size
    # the wrong target sector gives exactly zero
    other = subs[ks.index((2, 0))]
    assert me.ED_bracket(me.MBS64Vector(_rand(other.length, 5), other), op,
                         me.MBS64Vector(b, ket_sp)) == 0


def test_bracket_two_operators(gpu_ctx, oracle_mod):
    """ED_bracket(bra, op2, op1, ket) on the device against the oracle's restatement (itself
    pinned against dense products): density operators moving momentum between sectors, and
    a two-body upper-Hermitian op2 with the bra spanning all N-particle states."""
    import test_oracle as T
    for bra_b, bra, op2, op1, ket_b, ket, ref in T._bracket2_cases(oracle_mod):
        print(f"[bracket2] {len(op2.scats)} x {len(op1.scats)} terms, D {len(bra_b)} / "
              f"{len(ket_b)}", flush=True)
        bsp = me.HilbertSubspace(bra_b, bits=12)
        ksp = me.HilbertSubspace(ket_b, bits=12)
        got = me.ED_bracket(me.MBS64Vector(bra, bsp), op2, op1, me.MBS64Vector(ket, ksp))
        orc = oracle_mod.bracket2(bra_b, bra, _oracle_full(oracle_mod, op2),
                                  _oracle_full(oracle_mod, op1), ket_b, ket)
        assert abs(got - orc) <= 1e-12 * max(1.0, abs(orc))
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
        assert me.ED_bracket_threaded(me.MBS64Vector(bra, bsp), op2, op1,
                                      me.MBS64Vector(ket, ksp)) == got


def test_rdm1_closed_forms_and_oracle(gpu_ctx, oracle_mod):
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][0]
    bits = sp.bits
    # a basis state: occupations on the diagonal, nothing else
    e = np.zeros(sp.length, dtype=complex)
    e[17] = 1.0
    rho = me.PES_1rdm(me.MBS64Vector(e, sp))
    occ = [(int(sp.list[17]) >> k) & 1 for k in range(bits)]
    np.testing.assert_array_equal(rho, np.diag(np.array(occ, dtype=complex)))
    # an eigenvector: Hermitian, trace N, block-diagonal in momentum, equal to the oracle
    E, V = me.EDsolve(sp, op, N=1, method="map", method_info=False)
    psi = V[0].vec
    rho = me.PES_1rdm(V[0])
    np.testing.assert_allclose(rho, rho.conj().T, atol=1e-14)
    assert abs(np.trace(rho) - 5) < 1e-12
    for i in range(bits):
        for j in range(bits):
            one = oracle_mod.TermList(np.array([1.0 + 0j]), np.array([1 << i], dtype=np.uint64),
                                      np.array([1 << j], dtype=np.uint64))
            full = oracle_mod.full_list(one, False)
            ref = np.vdot(psi, oracle_mod.matvec(sp.list, full, psi))
            assert abs(rho[i, j] - ref) <= 1e-13, (i, j)


def test_rdm1_two_component_sector(gpu_ctx, oracle_mod):
    para = W.batched_para(1.2)
    op = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sp = me.ED_momentum_subspaces(para, (2, 3))[0][3]
    E, V = me.EDsolve(sp, op, N=1, method="map", method_info=False)
    psi = V[0].vec
    rho = me.PES_1rdm(V[0])
    for (i, j) in [(0, 0), (5, 5), (13, 13), (2, 14), (14, 2), (1, 1 + 12), (23, 11)]:
        one = oracle_mod.TermList(np.array([1.0 + 0j]), np.array([1 << i], dtype=np.uint64),
                                  np.array([1 << j], dtype=np.uint64))
        ref = np.vdot(psi, oracle_mod.matvec(sp.list, oracle_mod.full_list(one, False), psi))
        assert abs(rho[i, j] - ref) <= 1e-13, (i, j)
    assert abs(np.trace(rho) - 5) < 1e-12


def test_step_inner_prod(gpu_ctx):
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][2]
    E, V = me.EDsolve(sp, op, N=3, method="map", method_info=False)
    ones = np.ones(sp.bits, dtype=complex)
    got = me.ED_step_inner_prod(V[0], V[1], ones)
    assert abs(got - np.vdot(V[0].vec, V[1].vec)) < 1e-14
    theta = np.linspace(0.1, 1.7, sp.bits)
    u = np.exp(1j * theta)
    m = me.ED_step_inner_prod(V, V, u)
    phase = np.array([np.exp(1j * sum(theta[k] for k in range(sp.bits) if (int(s) >> k) & 1))
                      for s in sp.list])
    ref = np.array([[np.vdot(a.vec, b.vec * phase) for b in V] for a in V])
    np.testing.assert_allclose(m, ref, atol=1e-13)


def _bilayer_para(g, tunneling=False):
    import llt
    V = llt.LandauInteraction(llt.ReciprocalLattice.named(g["lattice"]),
                              *[tuple(c) for c in g["components"]])
    V.D_l, V.d_l, V.mix = g["D_l"], g["d_l"], g["mix"]
    V.V_intra, V.V_inter = list(g["V_intra"]), list(g["V_inter"])
    k = np.array(g["k_list"], dtype=np.int64)
    if not tunneling:
        return me.EDPara(k_list=k, Gk=tuple(g["Gk"]), Nc_mix=1, Nc_conserve=2, H_two=V), V

    def tunnel(kk, cf, ci, t_W0=0.0):      # the notebook's tunneling(k, cf, ci; t_W0)
        return t_W0 if cf != ci else 0.0
    return me.EDPara(k_list=k, Gk=tuple(g["Gk"]), Nc_mix=2, Nc_conserve=1, H_one=tunnel, H_two=V,
                     one_momentum_coordinate=False), V


@pytest.mark.parametrize("tunneling", [False, True])
def test_bilayer_halperin_energies_and_chern_number(gpu_ctx, tunneling):
    """End to end against 2.BilayerFQH_Halperin.ipynb: the ten lowest energies of sector (0,0)
    (device stored-matrix EDsolve) and the many-body Chern number of the two-fold ground
    space from a kshift plaquette (the three shifted Hamiltonians solved in one batch, the
    step overlaps on the device)."""
    import json
    import os
    from fractions import Fraction
    import llt
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "reference_printed_values.json")))["bilayer_halperin"]
    print(f"[bilayer] tunneling={tunneling}", flush=True)
    para, Vint = _bilayer_para(g, tunneling)
    N = 6 if tunneling else (3, 3)
    subs, ks = me.ED_momentum_subspaces(para, N)
    sp = subs[ks.index(tuple(g["sector"]))]
    one = [me.ED_scatterlist_onebody(para, t_W0=g["t_W0"])] if tunneling else []
    H = me.MBOperator(*one, me.ED_scatterlist_twobody(para), upper_hermitian=True)
    E, V = me.EDsolve(sp, H, N=10, method="sparse", method_info=False)
    ref = g["energies_per_Nk_tunneling" if tunneling else "energies_per_Nk_no_tunneling"]
    np.testing.assert_allclose(E / para.Nk, ref, rtol=1e-10, atol=1e-13)
    # kshift plaquette
    n = g["N_kshift"]
    f = Fraction(1, n)
    path = [(Fraction(0), Fraction(0)), (f, Fraction(0)), (f, f), (Fraction(0), f),
            (Fraction(0), Fraction(0))]
    ops = [me.MBOperator(*one, me.ED_scatterlist_twobody(para, kshift=path[i]),
                         upper_hermitian=True) for i in (1, 2, 3)]
    out = me.EDsolve_batch([(sp, o) for o in ops], N=6)
    vn = slice(g["vn_range"][0] - 1, g["vn_range"][1])
    psi = [V[vn]] + [vecs[vn] for _, vecs in out] + [V[vn]]
    W = [me.ED_step_inner_prod(psi[i + 1], psi[i],
                               llt.orbital_overlap(g["k_list"], g["Gk"], Vint, path[i + 1],
                                                   path[i]))
         for i in range(4)]
    P = W[0] @ W[1] @ W[2] @ W[3]
    chern = np.angle(np.linalg.det(P)) / (2 * np.pi) * n * n / len(psi[0])
    ref_c = g["chern_tunneling" if tunneling else "chern_no_tunneling"]
    assert abs(chern - ref_c) < 1e-8, (chern, ref_c)


def test_sphere_laughlin_unique_zero_mode_all_sectors(gpu_ctx, oracle_mod):
    """The reference's sphere example (19 orbitals, 7 electrons, one-dimensional angular
    momentum without modulo) with V_1 only: all 85 L_z sectors solved as one batch; the 1/3
    Laughlin state is the single zero-energy state, in the central sector L_z = 63, and the
    central sector's lowest 5 levels match the oracle's dense matrix."""
    import test_oracle as T
    para = T._sphere_para(19, [0.0, 1.0, 0.0, 0.0])
    subs, ks = me.ED_momentum_subspaces(para, 7)
    assert len(subs) == 85 and ks[0] == (21,) and ks[-1] == (105,)
    H = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    print(f"[sphere] {len(H.scats)} terms, 85 sectors, D max {max(sp.length for sp in subs)}",
          flush=True)
    res = me.EDsolve_batch([(sp, H) for sp in subs], N=1, want_vectors=False)
    e0 = np.array([E[0] for E, _ in res])
    zero = [ks[i][0] for i in np.flatnonzero(np.abs(e0) < 1e-10)]
    assert zero == [63]
    assert np.delete(e0, ks.index((63,))).min() > 1e-4
    sp = subs[ks.index((63,))]
    E, V = me.EDsolve(sp, H, N=5, method_info=False)
    ev = np.linalg.eigvalsh(oracle_mod.dense_matrix(sp.list, _oracle_full(oracle_mod, H)))[:5]
    np.testing.assert_allclose(E, ev, rtol=1e-10, atol=1e-12)
    assert abs(E[0]) < 1e-10 and E[1] > 1e-3


def test_pes_block_rdm_oracle(gpu_ctx, oracle_mod):
    """PES_MomtBlock_rdm on the device against the oracle's reference-order loops
    (particle_reduced_density_matrix.jl:37-88) within 1e-13: every 2-particle block of a
    complex state on the 3x5 mesh; a part holding all 5 particles (psi psi^H on psi's own
    sector, zero elsewhere); and a two-component sector split (1, 1) | (1, 2)."""
    w = W.small()
    subs, ks = me.ED_momentum_subspaces(w.para, 5)
    sp = subs[ks.index((1, 2))]
    psi = me.MBS64Vector(_rand(sp.length, 11), sp)
    parts, pks = me.ED_momentum_subspaces(w.para, 2)
    assert len(parts) == 15
    for A, k in zip(parts, pks):
        got = me.PES_MomtBlock_rdm(w.para, psi, A)
        ref = oracle_mod.pes_block_rdm(sp.list, psi.vec, A.list)
        assert np.abs(ref).max() > 1e-3
        assert np.abs(got - ref).max() <= 1e-13, k
    whole = me.PES_MomtBlock_rdm(w.para, psi, sp)
    assert np.abs(whole - np.outer(psi.vec, psi.vec.conj())).max() <= 1e-14
    assert np.abs(me.PES_MomtBlock_rdm(w.para, psi, subs[ks.index((0, 0))])).max() == 0.0
    para = W.batched_para(1.2)
    subs2, ks2 = me.ED_momentum_subspaces(para, (2, 3))
    sp2 = subs2[3]
    psi2 = me.MBS64Vector(_rand(sp2.length, 12), sp2)
    for A in me.ED_momentum_subspaces(para, (1, 1))[0][:5]:
        got = me.PES_MomtBlock_rdm(para, psi2, A)
        ref = oracle_mod.pes_block_rdm(sp2.list, psi2.vec, A.list)
        assert np.abs(ref).max() > 1e-3
        assert np.abs(got - ref).max() <= 1e-13


def test_pes_block_rdm_sphere_laughlin_counting(gpu_ctx):
    """The reference's sphere example (19 orbitals, 7 electrons, V_1 only): the particle
    entanglement spectrum of the Laughlin zero mode for a part of 3 particles. In each of the
    49 L_z blocks of part A the number of nonzero eigenvalues equals the number of
    (1,3)-admissible 3-particle configurations with that L_z (455 in all, C(15, 3)); the
    traces add up to C(7, 3) = 35."""
    import itertools
    import math
    import test_oracle as T
    para = T._sphere_para(19, [0.0, 1.0, 0.0, 0.0])
    H = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sp = me.ED_momentum_subspaces(para, 7, momentum_restriction=True,
                                  momentum_list=[(63,)])[0][0]
    E, V = me.EDsolve(sp, H, N=2, method_info=False)
    assert abs(E[0]) < 1e-10 < E[1]
    parts, pks = me.ED_momentum_subspaces(para, 3)
    assert len(parts) == 49
    total_trace, total_count = 0.0, 0
    for A, k in zip(parts, pks):
        rdm = me.PES_MomtBlock_rdm(para, V[0], A)
        assert np.abs(rdm - rdm.conj().T).max() <= 1e-14
        lam = np.linalg.eigvalsh(rdm)
        assert lam.min() > -1e-12
        count = int(np.sum(lam > 1e-10))
        want = sum(1 for c in itertools.combinations(range(19), 3)
                   if sum(c) == k[0] and c[1] - c[0] >= 3 and c[2] - c[1] >= 3)
        assert count == want, (k, count, want)
        total_count += count
        total_trace += float(np.trace(rdm).real)
    assert total_count == math.comb(15, 3) == 455
    assert abs(total_trace - 35) <= 1e-10


def test_pes_block_rdm_particle_hole_oracle(gpu_ctx, oracle_mod):
    """ph_transform on the device against the oracle within 1e-13: the single component of
    the 3x5 mesh flipped (5 particles -> 10 holes, parts of 2 holes), and the second of two
    conserved components on the 3x4 mesh flipped, where the complements of one sector spread
    over several momentum blocks."""
    w = W.small()
    subs, ks = me.ED_momentum_subspaces(w.para, 5)
    sp = subs[ks.index((2, 1))]
    psi = me.MBS64Vector(_rand(sp.length, 21), sp)
    ph = (1 << 15) - 1
    for A in me.ED_momentum_subspaces(w.para, 2)[0][:6]:
        got = me.PES_MomtBlock_rdm(w.para, psi, A, [True])
        ref = oracle_mod.pes_block_rdm(sp.list, psi.vec, A.list, ph)
        assert np.abs(ref).max() > 1e-3
        assert np.abs(got - ref).max() <= 1e-13
    para = W.batched_para(0.8)
    sp2 = me.ED_momentum_subspaces(para, (2, 3))[0][5]
    psi2 = me.MBS64Vector(_rand(sp2.length, 22), sp2)
    ph2 = ((1 << 12) - 1) << 12
    trace = 0.0
    for A in me.ED_momentum_subspaces(para, (1, 1))[0]:
        got = me.PES_MomtBlock_rdm(para, psi2, A, [False, True])
        ref = oracle_mod.pes_block_rdm(sp2.list, psi2.vec, A.list, ph2)
        assert np.abs(ref).max() > 1e-3
        assert np.abs(got - ref).max() <= 1e-13
        trace += float(np.trace(got).real)
    assert abs(trace - 2 * 9) <= 1e-11
