"""GPU parity: libedm.so kernels against the CPU oracle, through the C-ABI.

Bars: basis states and (row, column, term, sign) connectivity bit-exact; y = H x within
1e-13 of max|y| (the GPU sums a row's terms in a different order than the reference);
eigenvalues within 1e-10 relative of dense LAPACK / ARPACK on the oracle's matrix and of the
reference's printed values; eigenvector overlap > 1 - 1e-9 for nondegenerate levels.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import momentumed as me
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import check

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_printed_values.json")))
Y_TOL = 1e-13


def _system(para, N_each):
    return dev.system_struct(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, N_each)


def _oracle_full(O, op: me.MBOperator):
    amp, out, inn = op.arrays()
    return O.full_list(O.TermList(amp, out, inn), op.upper_hermitian)


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _gpu_matvec(ctx, dop, basis, x):
    y = np.empty_like(x)
    check(ctx.L.edm_matvec_host(ctx.h, dop.h, basis.h, x.ctypes.data, y.ctypes.data), ctx.h)
    return y


def _gpu_matvec_block(ctx, dop, basis, xs):
    """edm_matvec_block on device copies of the host vectors xs; returns the host results."""
    import torch
    X = [torch.from_numpy(xv).cuda() for xv in xs]
    Y = [torch.empty_like(t) for t in X]
    work = torch.empty(len(xs) * basis.D, dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    xp = (ctypes.c_void_p * len(xs))(*[t.data_ptr() for t in X])
    yp = (ctypes.c_void_p * len(xs))(*[t.data_ptr() for t in Y])
    check(ctx.L.edm_matvec_block(ctx.h, dop.h, basis.h, len(xs), xp, yp, work.data_ptr()), ctx.h)
    ctx.synchronize()
    return [t.cpu().numpy() for t in Y]


def _gpu_connectivity(ctx, dop, basis, row0, nrows):
    counts = np.zeros(max(nrows, 1), dtype=np.int64)
    check(ctx.L.edm_connectivity_counts(ctx.h, dop.h, basis.h, row0, nrows, counts.ctypes.data), ctx.h)
    rowptr = np.zeros(nrows + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(counts[:nrows])
    nnz = int(rowptr[-1])
    cols = np.zeros(max(nnz, 1), dtype=np.int64)
    terms = np.zeros(max(nnz, 1), dtype=np.int32)
    negs = np.zeros(max(nnz, 1), dtype=np.int8)
    check(ctx.L.edm_connectivity(ctx.h, dop.h, basis.h, row0, nrows, rowptr.ctypes.data,
                                 cols.ctypes.data, terms.ctypes.data, negs.ctypes.data), ctx.h)
    return rowptr, cols[:nnz], terms[:nnz], negs[:nnz]


def _sorted_rows(rowptr, cols, terms, negs):
    out = []
    for r in range(len(rowptr) - 1):
        a, b = rowptr[r], rowptr[r + 1]
        out.append(sorted(zip(cols[a:b].tolist(), terms[a:b].tolist(), negs[a:b].tolist())))
    return out


# ------------------------------------------------------------------ enumeration (a1)
@pytest.mark.parametrize("entry", GOLDEN["sector_sizes"], ids=lambda e: e["source"][-30:])
def test_enumeration_matches_oracle_and_reference(gpu_ctx, oracle_mod, entry):
    print(f"[enum] {entry.get('mesh', 'k_list')} N={entry['N']} Nc_mix={entry['Nc_mix']} "
          f"Ncc={entry['Nc_conserve']}", flush=True)
    k = np.array(entry["k_list"], dtype=np.int64) if "k_list" in entry else W.mesh(*entry["mesh"])
    para = me.EDPara(k_list=k, Gk=tuple(entry["Gk"]), Nc_mix=entry["Nc_mix"],
                     Nc_conserve=entry["Nc_conserve"])
    sysd = _system(para, entry["N"])
    nsec = ctypes.c_int64()
    check(gpu_ctx.L.edm_num_sectors(ctypes.byref(sysd), ctypes.byref(nsec)))
    counts = np.zeros(nsec.value, dtype=np.int64)
    check(gpu_ctx.L.edm_sector_sizes(gpu_ctx.h, ctypes.byref(sysd), counts.ctypes.data), gpu_ctx.h)
    lists, ks = oracle_mod.enumerate_sectors(k, entry["Gk"], entry["Nc_mix"],
                                             entry["Nc_conserve"], entry["N"])
    assert [c for c in counts if c] == [len(v) for v in lists]
    if "sizes" in entry:
        assert [c for c in counts if c] == entry["sizes"]
    if "momenta" in entry:     # the public API, all sectors in one pass
        subs, sks = me.ED_momentum_subspaces(para, entry["N"])
        assert [list(q) for q in sks] == entry["momenta"]
        assert [sp.length for sp in subs] == entry["sizes"]
    j = 0
    alls = dev.DeviceBasis.enumerate_all(gpu_ctx, sysd)
    assert len(alls) == nsec.value
    for s in range(nsec.value):
        b = dev.DeviceBasis.enumerate(gpu_ctx, sysd, s)
        assert b.D == counts[s] == alls[s].D
        assert b.rank_mode == alls[s].rank_mode
        if b.D:
            np.testing.assert_array_equal(b.download(), lists[j])
            np.testing.assert_array_equal(alls[s].download(), lists[j])
            j += 1


@pytest.mark.parametrize("case", ["free_momentum_4x4_N3", "medium_3x9_N9"])
def test_enumerate_all_sectors_in_one_pass(gpu_ctx, oracle_mod, case):
    """edm_basis_enumerate_all against the oracle and the per-sector enumeration: 100
    sectors with Gk = (0, 0) (two counting chunks of 64, seven writing chunks of 16, empty
    sectors among them) and the 27 sectors of the 3x9 N=9 system."""
    if case.startswith("free"):
        k, gk, N = W.mesh(4, 4), (0, 0), 3
    else:
        k, gk, N = W.mesh(3, 9), (3, 9), 9
    para = me.EDPara(k_list=k, Gk=gk)
    sysd = _system(para, [N])
    alls = dev.DeviceBasis.enumerate_all(gpu_ctx, sysd)
    lists, ks = oracle_mod.enumerate_sectors(k, gk, 1, 1, [N])
    mlist = me.preparation._momentum_list(para, [N])
    assert len(alls) == len(mlist)
    nonempty = [(mlist[s], b) for s, b in enumerate(alls) if b.D]
    assert [kk for kk, _ in nonempty] == list(ks)
    for (kk, b), ref in zip(nonempty, lists):
        np.testing.assert_array_equal(b.download(), ref)
    for s in (0, len(alls) // 2, len(alls) - 1):
        one = dev.DeviceBasis.enumerate(gpu_ctx, sysd, s)
        assert one.D == alls[s].D and one.rank_mode == alls[s].rank_mode
        if one.D:
            np.testing.assert_array_equal(one.download(), alls[s].download())
    subs, sks = me.ED_momentum_subspaces(para, N)
    assert sks == list(ks)
    assert [sp.length for sp in subs] == [len(v) for v in lists]


def test_api_momentum_subspaces(gpu_ctx, oracle_mod):
    w = W.small()
    subspaces, ks = me.ED_momentum_subspaces(w.para, 5)
    lists, oks = oracle_mod.enumerate_sectors(w.para.k_list, w.para.Gk, 1, 1, [5])
    assert ks == oks
    for sp, ref in zip(subspaces, lists):
        np.testing.assert_array_equal(sp.list, ref)
    sub2, ks2 = me.ED_momentum_subspaces(w.para, 5, momentum_restriction=True,
                                         momentum_list=[(1, 1), (4, 5)])
    assert ks2 == [(1, 1), (1, 0)]


def test_rank_tables_modes(gpu_ctx, oracle_mod):
    w = W.medium()
    sysd = _system(w.para, w.N_each)
    b = dev.DeviceBasis.enumerate(gpu_ctx, sysd, 0)
    assert b.D == 173613 and b.rank_mode == 1
    full = b.download()
    assert dev.DeviceBasis.upload(gpu_ctx, full, 27, sysd, 0).rank_mode == 1
    # without the system the tables are derived from the states (mode 2), also for a partial
    # list that splits into equal blocks (every other state); drop_rank returns to search
    b2 = dev.DeviceBasis.upload(gpu_ctx, full, 27)
    assert b2.rank_mode == 2
    assert dev.DeviceBasis.upload(gpu_ctx, full[::2], 27, sysd, 0).rank_mode == 2
    b2.drop_rank()
    assert b2.rank_mode == 0


# ------------------------------------------------------------------ connectivity (a6-a8)
def test_full_list_matches_oracle(gpu_ctx, oracle_mod):
    for w in (W.small(), W.medium()):
        op = w.operator()
        dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
        amp, out, inn, par = dop.full_list()
        ref = _oracle_full(oracle_mod, op)
        np.testing.assert_array_equal(out, ref.out)
        np.testing.assert_array_equal(inn, ref.inn)
        np.testing.assert_array_equal(par, ref.parity)
        np.testing.assert_array_equal(amp, ref.amp)


def test_connectivity_bit_exact_small_all_sectors(gpu_ctx, oracle_mod):
    w = W.small()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    full = _oracle_full(oracle_mod, op)
    sysd = _system(w.para, w.N_each)
    lists, _ = oracle_mod.enumerate_sectors(w.para.k_list, w.para.Gk, 1, 1, [5])
    for s, ref_basis in enumerate(lists):
        b = dev.DeviceBasis.enumerate(gpu_ctx, sysd, s)
        assert b.rank_mode == 1
        g = _sorted_rows(*_gpu_connectivity(gpu_ctx, dop, b, 0, b.D))
        o = _sorted_rows(*oracle_mod.connectivity(ref_basis, full))
        assert g == o, f"sector {s}"


def test_connectivity_bit_exact_medium_rows(gpu_ctx, oracle_mod):
    w = W.medium()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    full = _oracle_full(oracle_mod, op)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)
    basis = b.download()
    for row0, n in ((0, 3000), (b.D // 2, 3000), (b.D - 3000, 3000)):
        g = _sorted_rows(*_gpu_connectivity(gpu_ctx, dop, b, row0, n))
        o = _sorted_rows(*oracle_mod.connectivity(basis, full, row0, n))
        assert g == o
    # every row's count (the nonzero structure of the whole sector)
    counts = np.zeros(b.D, dtype=np.int64)
    check(gpu_ctx.L.edm_connectivity_counts(gpu_ctx.h, dop.h, b.h, 0, b.D, counts.ctypes.data), gpu_ctx.h)
    rp, _, _, _ = oracle_mod.connectivity(basis, full, 0, b.D)
    np.testing.assert_array_equal(counts, np.diff(rp))


# ------------------------------------------------------------------ matvec (a6)
@pytest.mark.parametrize("name", ["small", "medium"])
def test_matvec_matches_oracle(gpu_ctx, oracle_mod, name):
    w = W.get(name)
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    full = _oracle_full(oracle_mod, op)
    sysd = _system(w.para, w.N_each)
    nsec = 15 if name == "small" else 1
    for s in range(nsec):
        b = dev.DeviceBasis.enumerate(gpu_ctx, sysd, s)
        basis = b.download()
        x = _rand(b.D, 100 + s)
        y = _gpu_matvec(gpu_ctx, dop, b, x)
        yo = oracle_mod.matvec(basis, full, x)
        assert np.abs(y - yo).max() <= Y_TOL * np.abs(yo).max()
        # adjoint of a Hermitian operator is itself
        lm = dev.ROCmLinearMap(dop, b).adjoint()
        ya = lm(x)
        assert np.abs(ya - yo).max() <= Y_TOL * np.abs(yo).max()
        # rank tables derived from the state list (the reference's call sequence: a bare
        # list, no system) and binary-search lookup give the identical result bit for bit
        b2 = dev.DeviceBasis.upload(gpu_ctx, basis, w.para.bits)
        assert b2.rank_mode == 2
        np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, dop, b2, x), y)
        b2.drop_rank()
        assert b2.rank_mode == 0
        np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, dop, b2, x), y)


@pytest.mark.parametrize("variant", ["basic", "lo", "hi", "u1", "u2", "u4", "u1a"])
@pytest.mark.parametrize("swizzle", ["0", "1"])
def test_matvec_kernel_variants(gpu_ctx, oracle_mod, monkeypatch, variant, swizzle):
    """Every bucketed-kernel variant (global table; LDS table with low- or high-bit-first pair
    order; with and without the XCD block remap) against the oracle."""
    monkeypatch.setenv("EDM_MATVEC", variant)
    monkeypatch.setenv("EDM_XCD_SWIZZLE", swizzle)
    w = W.medium()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)
    x = _rand(b.D, 31)
    yo = oracle_mod.matvec(b.download(), _oracle_full(oracle_mod, op), x)
    assert np.abs(_gpu_matvec(gpu_ctx, dop, b, x) - yo).max() <= Y_TOL * np.abs(yo).max()


def test_matvec_tune_keeps_y(gpu_ctx, oracle_mod):
    """The autotuner only changes the workgroup size, where the amplitudes are read from and
    how many lanes share a row: y is bitwise the same before and after, for whichever kernel
    it picks, and the LDS-amplitude kernel gives the same y as the global-table one."""
    w = W.medium()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)
    H = dev.ROCmLinearMap(dop, b)
    x = _rand(b.D, 77)
    y0 = _gpu_matvec(gpu_ctx, dop, b, x)
    blk, ms = H.tune()
    print(f"[tune] medium: chose {blk:#x}, ms {ms}", flush=True)
    assert blk & 0xFFFF in (256, 512, 1024) and all(v > 0 for v in ms.values())
    # the medium operator's amplitudes fit in LDS, and so does its 13-bit low rank table
    assert len(ms) == 12
    assert H.kernel() == blk
    np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, dop, b, x), y0)
    import os
    os.environ["EDM_MATVEC"] = "u1a"
    try:
        np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, dop, b, x), y0)
    finally:
        del os.environ["EDM_MATVEC"]
    part = dev.DeviceBasis.upload(gpu_ctx, b.download()[::5].copy(), w.para.bits)
    Hp = dev.ROCmLinearMap(dop, part)
    xp = _rand(part.D, 78)
    yp0 = _gpu_matvec(gpu_ctx, dop, part, xp)
    blk2, _ = Hp.tune()
    assert blk2 & 0xFFFF in (256, 512, 1024)
    np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, dop, part, xp), yp0)


def test_matvec_split_rows_bitwise(gpu_ctx, oracle_mod, monkeypatch):
    """The split-row kernel (2, 4, 8, 16 lanes per row, the row's first lane adding the
    parked contributions in entry order) gives bitwise the one-lane y: a complete sector with
    rank tables, a partial basis (binary search, partners missing), a system with one-body
    terms, and two conserved components; the one-lane y matches the oracle."""
    cases = []
    w = W.medium()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, w.operator(), w.para.bits)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)
    cases.append(("medium", w.operator(), dop, b))
    part = dev.DeviceBasis.upload(gpu_ctx, b.download()[::7].copy(), w.para.bits)
    cases.append(("medium/7", w.operator(), dop, part))
    ws = W.small()
    dops = dev.DeviceOperator.from_mboperator(gpu_ctx, ws.operator(), ws.para.bits)
    for sec in (0, 7):
        cases.append((f"small s{sec}", ws.operator(), dops,
                      dev.DeviceBasis.enumerate(gpu_ctx, _system(ws.para, ws.N_each), sec)))
    para = W.batched_para(1.3)
    op2 = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    dop2 = dev.DeviceOperator.from_mboperator(gpu_ctx, op2, para.bits)
    cases.append(("two-component", op2, dop2,
                  dev.DeviceBasis.enumerate(gpu_ctx, _system(para, (2, 3)), 3)))
    # 4x6 mesh per component: a 57 KB term image, so image + exchange slots pass 64 KiB and
    # the kernel runs with the raised LDS limit
    k46 = W.mesh(4, 6)
    para46 = me.EDPara(k_list=k46, Gk=(4, 6), Nc_conserve=2,
                       H_two=W.CoulombGaussian(k46, (4, 6), c=0.25, inter_ratio=0.8),
                       two_momentum_coordinate=False)
    op46 = me.MBOperator(me.ED_scatterlist_twobody(para46), upper_hermitian=True)
    dop46 = dev.DeviceOperator.from_mboperator(gpu_ctx, op46, para46.bits)
    cases.append(("4x6 two-component", op46, dop46,
                  dev.DeviceBasis.enumerate(gpu_ctx, _system(para46, (2, 1)), 5)))
    for name, op, d, basis in cases:
        x = _rand(basis.D, 91)
        monkeypatch.setenv("EDM_MATVEC_LANES", "1")
        y1 = _gpu_matvec(gpu_ctx, d, basis, x)
        if basis.D <= 20000:
            yo = oracle_mod.matvec(basis.download(), _oracle_full(oracle_mod, op), x)
            assert np.abs(y1 - yo).max() <= Y_TOL * np.abs(yo).max(), name
        for lanes in ("2", "4", "8", "16"):
            monkeypatch.setenv("EDM_MATVEC_LANES", lanes)
            np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, d, basis, x), y1,
                                          err_msg=f"{name} lanes={lanes}")
        print(f"[split] {name}: D={basis.D} bitwise equal for 2..16 lanes", flush=True)


def test_matvec_lds_rank_table_bitwise(gpu_ctx, oracle_mod, monkeypatch):
    """The kernel that stages the low-half rank table in LDS (k_matvec_narrow) walks each row
    as the global-table kernel does and looks partners up through the same kind of split:
    its y is bitwise the one-lane kernel's (1024- and 512-row workgroups, with the walk that
    gathers x into LDS slots while it goes on and with the plain walk), and the one-lane y is
    the oracle's to 1e-13. Cases: rank tables whose low half is already at most 14 bits
    (3x5 N=5, 3x9 N=9, two conserved components on 3x4 meshes), used as they are, and a
    30-orbital sector (5x6 mesh, N=7; 15 + 15 bit tables) whose 14 + 16 bit split the library
    derives from the state list on the device and checks on every state; and a general
    (non-Hermitian) operator with complex diagonal terms on a bare list of states."""
    cases = []
    ws = W.small()
    dops = dev.DeviceOperator.from_mboperator(gpu_ctx, ws.operator(), ws.para.bits)
    cases.append(("small s3", ws.operator(), dops,
                  dev.DeviceBasis.enumerate(gpu_ctx, _system(ws.para, ws.N_each), 3)))
    w = W.medium()
    dopm = dev.DeviceOperator.from_mboperator(gpu_ctx, w.operator(), w.para.bits)
    cases.append(("medium", w.operator(), dopm,
                  dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)))
    para = W.batched_para(0.7)
    op2 = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    dop2 = dev.DeviceOperator.from_mboperator(gpu_ctx, op2, para.bits)
    cases.append(("two-component", op2, dop2,
                  dev.DeviceBasis.enumerate(gpu_ctx, _system(para, (3, 4)), 5)))
    k56 = W.mesh(5, 6)
    para56 = me.EDPara(k_list=k56, Gk=(5, 6), H_two=W.CoulombGaussian(k56, (5, 6), c=0.25),
                       two_momentum_coordinate=False)
    op56 = me.MBOperator(me.ED_scatterlist_twobody(para56), upper_hermitian=True)
    dop56 = dev.DeviceOperator.from_mboperator(gpu_ctx, op56, para56.bits)
    b56 = dev.DeviceBasis.enumerate(gpu_ctx, _system(para56, (7,)), 0)
    assert b56.rank_mode == 1 and b56.D > 50000
    cases.append(("5x6 N=7", op56, dop56, b56))
    # a general operator (not Hermitian) with complex diagonal terms on all 6-particle states
    # of 16 orbitals: the pipelined walk sums the diagonal apart from the walk
    rng = np.random.default_rng(41)
    terms = list(_random_general_op(rng, 16).scats)
    for _ in range(12):
        a, b_ = rng.choice(16, 2, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a), int(b_), int(b_),
                                int(a), bits=16))
    for a in range(1, 17, 3):
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), a, a, bits=16))
    opg = me.MBOperator(terms, upper_hermitian=False)
    states = np.array([v for v in range(1 << 16) if bin(v).count("1") == 6], dtype=np.uint64)
    bg = dev.device_basis_of(me.HilbertSubspace(states, bits=16), gpu_ctx)
    cases.append(("general, complex diagonal", opg,
                  dev.DeviceOperator.from_mboperator(gpu_ctx, opg, 16), bg))
    for name, op, d, basis in cases:
        x = _rand(basis.D, 93)
        H = dev.ROCmLinearMap(d, basis)
        monkeypatch.setenv("EDM_MATVEC", "u1")
        monkeypatch.setenv("EDM_MATVEC_LANES", "1")
        assert H.kernel() & 0xFFFF == 512 and not H.kernel() & dev.TUNE_NARROW
        y1 = _gpu_matvec(gpu_ctx, d, basis, x)
        monkeypatch.delenv("EDM_MATVEC_LANES")
        rows = np.arange(basis.D) if basis.D <= 20000 else \
            np.random.default_rng(3).integers(0, basis.D, 3000)
        yo = oracle_mod.matvec(basis.download(), _oracle_full(oracle_mod, op), x, rows=rows)
        assert np.abs(y1[rows] - yo).max() <= Y_TOL * np.abs(yo).max(), name
        for v, blk in (("u1n", 1024), ("u1n512", 512)):
            monkeypatch.setenv("EDM_MATVEC", v)
            assert H.kernel() == blk | dev.TUNE_NARROW, (name, v, hex(H.kernel()))
            # pipelined walk where it keeps the occupancy (default), plain walk, pipelined
            # walk wherever it fits
            for pipe in (None, "0", "2"):
                if pipe is None:
                    monkeypatch.delenv("EDM_NARROW_PIPE", raising=False)
                else:
                    monkeypatch.setenv("EDM_NARROW_PIPE", pipe)
                np.testing.assert_array_equal(_gpu_matvec(gpu_ctx, d, basis, x), y1,
                                              err_msg=f"{name} {v} pipe {pipe}")
            monkeypatch.delenv("EDM_NARROW_PIPE")
        # the block kernel with the same lookup: every column bitwise the single matvec
        monkeypatch.setenv("EDM_MATVEC", "u1")
        nv = 4
        xs = [np.ascontiguousarray(x * (v + 1) + _rand(basis.D, 94 + v)) for v in range(nv)]
        ys = [_gpu_matvec(gpu_ctx, d, basis, xv) for xv in xs]
        for wg in ("512", "1024"):
            monkeypatch.setenv("EDM_BLOCK_NARROW", wg)
            yb = _gpu_matvec_block(gpu_ctx, d, basis, xs)
            for v in range(nv):
                np.testing.assert_array_equal(yb[v], ys[v], err_msg=f"{name} block {wg} col {v}")
        monkeypatch.delenv("EDM_BLOCK_NARROW")
        print(f"[lds-rank] {name}: D={basis.D} bitwise equal (1024, 512, block)", flush=True)


@pytest.mark.parametrize("N", [0, 1, 2, 13, 14, 15])
def test_edge_fillings(gpu_ctx, oracle_mod, N):
    """Empty, single-particle, nearly full and full fillings of the 15-orbital 3x5 system
    (the reference's colex cases t = 0 and t = n): sectors and states bit-exact against the
    oracle (one vacuum sector for N = 0, one full state for N = 15), and each sector's
    lowest energy through EDsolve equal to the oracle's dense matrix within 1e-10, both with
    the default method (which switches to dense for these tiny sectors, as in the reference)
    and with the matrix-free map (Krylov spaces down to a single vector)."""
    w = W.small()
    op = w.operator()
    subs, ks = me.ED_momentum_subspaces(w.para, N)
    o_lists, o_ks = oracle_mod.enumerate_sectors(w.para.k_list, w.para.Gk, 1, 1, [N])
    assert ks == o_ks and len(subs) == len(o_lists)
    full = _oracle_full(oracle_mod, op)
    for sp, ob in zip(subs, o_lists):
        np.testing.assert_array_equal(sp.list, ob)
        e_ref = np.linalg.eigvalsh(oracle_mod.dense_matrix(ob, full))[0]
        for method in ("sparse", "map"):   # sparse switches to dense below 100 states
            E, V = me.EDsolve(sp, op, N=1, method=method, method_info=False)
            assert abs(E[0] - e_ref) <= 1e-10 * max(1.0, abs(e_ref)), (method, sp.length)
            assert abs(np.linalg.norm(V[0].vec) - 1.0) < 1e-12
    if N == 0:
        assert [int(s) for s in subs[0].list] == [0] and abs(E[0]) < 1e-15


@pytest.mark.parametrize("layout", ["mixed_conserved", "three_conserved", "three_dimensional",
                                    "open_direction", "sixty_four_bits"])
def test_component_layouts(gpu_ctx, oracle_mod, layout):
    """Mixed components inside two conserved ones (with one-body tunnelling), three conserved
    components, momenta in three dimensions (2x2x3 mesh), a mesh periodic in x and open in y
    (Gk = (3, 0)), and 64 orbitals (two mixed components on 4x8; partners found by binary
    search, bit 63 in use): every sector's basis bit-exact
    against the oracle, Hx within 1e-13 and the lowest eigenvalues within 1e-10 of the
    oracle's dense matrix, for every sector (the oracle's Hamiltonian for all three is pinned
    against Jordan-Wigner)."""
    import test_oracle as T
    if layout == "mixed_conserved":
        para, N, ops = T._para_mixed_conserved(), (2, 2), True
    elif layout == "three_conserved":
        para, N, ops = T._para_three_conserved(), (1, 2, 1), False
    elif layout == "three_dimensional":
        para, N, ops = T._para_3d(), (4,), False
    elif layout == "open_direction":
        para, N, ops = T._para_open(), (4,), False
    else:
        para, N, ops = T._para_64(), (2,), True
    lists = [me.ED_scatterlist_twobody(para)]
    if ops:
        lists.insert(0, me.ED_scatterlist_onebody(para))
    op = me.MBOperator(*lists, upper_hermitian=True)
    subs, ks = me.ED_momentum_subspaces(para, N)
    o_lists, o_ks = oracle_mod.enumerate_sectors(para.k_list, para.Gk, para.Nc_mix,
                                                 para.Nc_conserve, list(N))
    assert ks == o_ks and len(subs) == len(o_lists)
    full = _oracle_full(oracle_mod, op)
    for sp, ob, k in zip(subs, o_lists, ks):
        np.testing.assert_array_equal(sp.list, ob)
        H = oracle_mod.dense_matrix(ob, full)
        x = _rand(sp.length, 5)
        dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, para.bits)
        y = _gpu_matvec(gpu_ctx, dop, dev.device_basis_of(sp, gpu_ctx), x)
        assert np.abs(y - H @ x).max() <= Y_TOL * np.abs(H @ x).max(), k
        if sp.length >= 4:
            E, _ = me.EDsolve(sp, op, N=3, method="map", method_info=False)
            np.testing.assert_allclose(E, np.linalg.eigvalsh(H)[:3], rtol=1e-10, atol=1e-12)
    print(f"[layout] {layout}: {len(subs)} sectors, D {[sp.length for sp in subs]}", flush=True)


@pytest.mark.parametrize("passes", ["2", "3"])
def test_matvec_windowed_passes(gpu_ctx, oracle_mod, monkeypatch, passes):
    """The matvec split into passes over windows of x (accumulating into y)."""
    monkeypatch.setenv("EDM_MATVEC", "u2")
    monkeypatch.setenv("EDM_MATVEC_SPLIT", passes)
    w = W.medium()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)
    x = _rand(b.D, 32)
    yo = oracle_mod.matvec(b.download(), _oracle_full(oracle_mod, op), x)
    assert np.abs(_gpu_matvec(gpu_ctx, dop, b, x) - yo).max() <= Y_TOL * np.abs(yo).max()


def test_matvec_partial_basis_drops_targets(gpu_ctx, oracle_mod):
    w = W.small()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    full = _oracle_full(oracle_mod, op)
    lists, _ = oracle_mod.enumerate_sectors(w.para.k_list, w.para.Gk, 1, 1, [5])
    part = lists[0][::3].copy()
    b = dev.DeviceBasis.upload(gpu_ctx, part, w.para.bits, _system(w.para, w.N_each), 0)
    # the every-third list splits into equal blocks, so it gets derived rank tables, but
    # targets fall outside it: the closure check sends its lookups to binary search
    assert b.rank_mode == 2
    x = _rand(len(part), 7)
    yo = oracle_mod.matvec(part, full, x)
    assert np.abs(_gpu_matvec(gpu_ctx, dop, b, x) - yo).max() <= Y_TOL * np.abs(yo).max()


def test_matvec_non_hermitian_and_adjoint(gpu_ctx, oracle_mod):
    """A general (not upper-Hermitian) list: H x and H^dagger x against the oracle's dense H."""
    rng = np.random.default_rng(11)
    bits = 10
    terms = []
    for _ in range(60):
        a, b_, c, d = rng.choice(bits, 4, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a), int(b_), int(c), int(d), bits=bits))
    for _ in range(20):
        a, c = rng.choice(bits, 2, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a), int(c), bits=bits))
    op = me.MBOperator(terms, upper_hermitian=False)
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, bits)
    full = _oracle_full(oracle_mod, op)
    basis = np.array([s for s in range(1 << bits) if bin(s).count("1") == 4], dtype=np.uint64)
    H = oracle_mod.dense_matrix(basis, full)
    b = dev.DeviceBasis.upload(gpu_ctx, basis, bits)
    x = _rand(len(basis), 12)
    lm = dev.ROCmLinearMap(dop, b)
    y = lm(x)
    np.testing.assert_allclose(y, H @ x, rtol=0, atol=Y_TOL * np.abs(H @ x).max())
    ya = lm.adjoint()(x)
    np.testing.assert_allclose(ya, H.conj().T @ x, rtol=0, atol=Y_TOL * np.abs(ya).max())


def test_matvec_three_body_generic_path(gpu_ctx, oracle_mod):
    """Terms the bucketed table cannot hold (3-body) run in the generic kernel."""
    rng = np.random.default_rng(13)
    bits = 12
    terms = []
    for _ in range(40):
        o = rng.choice(bits, 3, replace=False) + 1
        i = rng.choice(bits, 3, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), *map(int, o), *map(int, i),
                                bits=bits, upper_hermitian=True))
    terms += [me.Scatter(1.5, 1, 2, bits=bits, upper_hermitian=True)]
    op = me.MBOperator(terms, upper_hermitian=True)
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, bits)
    full = _oracle_full(oracle_mod, op)
    basis = np.array([s for s in range(1 << bits) if bin(s).count("1") == 5], dtype=np.uint64)
    b = dev.DeviceBasis.upload(gpu_ctx, basis, bits)
    x = _rand(len(basis), 14)
    yo = oracle_mod.matvec(basis, full, x)
    assert np.abs(_gpu_matvec(gpu_ctx, dop, b, x) - yo).max() <= Y_TOL * np.abs(yo).max()


def test_matvec_device_pointers(gpu_ctx, oracle_mod):
    import torch
    w = W.small()
    op = w.operator()
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 3)
    x = _rand(b.D, 21)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    dev.ROCmLinearMap(dop, b)(yd, xd)
    yo = oracle_mod.matvec(b.download(), _oracle_full(oracle_mod, op), x)
    assert np.abs(yd.cpu().numpy() - yo).max() <= Y_TOL * np.abs(yo).max()


# ------------------------------------------------------------------ eigensolve (a9-a10)
def _check_vectors(H, vals, vecs, ref_vals, ref_vecs):
    for i in range(len(vals)):
        v = vecs[i]
        r = H @ v - vals[i] * v
        assert np.linalg.norm(r) < 1e-9
        gap = min([abs(ref_vals[i] - ref_vals[j]) for j in range(len(ref_vals)) if j != i] + [1.0])
        if gap > 1e-6:
            ov = abs(np.vdot(ref_vecs[:, i], v))
            assert ov > 1 - 1e-9


def test_eigsolve_small_all_sectors(gpu_ctx, oracle_mod):
    w = W.small()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    subspaces, ks = me.ED_momentum_subspaces(w.para, 5)
    for sp in subspaces:
        E, V = me.EDsolve(sp, op, N=w.nev, method="map", device="gpu", method_info=False)
        H = oracle_mod.dense_matrix(sp.list, full)
        ev, evec = np.linalg.eigh(H)
        np.testing.assert_allclose(E, ev[:w.nev], rtol=1e-10, atol=1e-12)
        _check_vectors(H, E, [v.vec for v in V], ev[: w.nev + 1], evec)


def test_eigsolve_reference_laughlin_energies(gpu_ctx):
    """End to end through the public API: the energies printed by the reference notebook."""
    import llt
    g = GOLDEN["energies_per_Nk"][0]
    V = llt.LandauInteraction(llt.ReciprocalLattice.named("triangular"), (1, 0, 1, 0))
    V.D_l, V.mix = g["D_l"], g["mix"]
    para = me.EDPara(k_list=W.mesh(*g["mesh"]), Gk=tuple(g["Gk"]), H_two=V)
    subspaces, ks = me.ED_momentum_subspaces(para, g["N"])
    H = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sp = subspaces[ks.index(tuple(g["sector"]))]
    E, _ = me.EDsolve(sp, H, N=10, method="map", device="gpu", method_info=False)
    np.testing.assert_allclose(E / para.Nk, g["values"], rtol=1e-10, atol=0)


def test_eigsolve_medium_against_arpack(gpu_ctx, oracle_mod):
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    w = W.medium()
    op = w.operator()
    full = _oracle_full(oracle_mod, op)
    subspaces, ks = me.ED_momentum_subspaces(w.para, 9, momentum_restriction=True,
                                             momentum_list=[(0, 0)])
    sp = subspaces[0]
    E, V = me.EDsolve(sp, op, N=w.nev, method="map", device="gpu", method_info=False,
                      krylovdim=w.krylovdim, tol=w.tol)
    basis = sp.list
    rp, cols, terms, negs = oracle_mod.connectivity(basis, full)
    vals = full.amp[terms] * np.where(negs, -1.0, 1.0)
    Hs = sps.csr_matrix((vals, cols, rp), shape=(len(basis), len(basis)))
    ref = np.sort(spl.eigsh(Hs, k=w.nev, which="SA", tol=1e-14, ncv=60)[0])
    np.testing.assert_allclose(E, ref, rtol=1e-10, atol=0)
    for i, v in enumerate(V):
        x = v.vec
        r = oracle_mod.matvec(basis, full, x) - E[i] * x
        assert np.linalg.norm(r) < 1e-9
    G = np.array([[np.vdot(a.vec, b.vec) for b in V] for a in V])
    assert np.abs(G - np.eye(len(V))).max() < 1e-10


def test_eigsolve_seeded_and_vec0(gpu_ctx):
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][0]
    H = dev.create_gpu_linearmap(op, sp)
    v1, _, i1 = dev.krylov_map_solve(H, 4, seed=7)
    v2, _, i2 = dev.krylov_map_solve(H, 4, seed=7)
    np.testing.assert_array_equal(v1, v2)          # deterministic for a fixed seed
    v3, _, _ = dev.krylov_map_solve(H, 4, vec0=np.ones(len(sp)) + 0.5j)
    np.testing.assert_allclose(v3, v1, rtol=1e-11, atol=1e-12)
    assert i1["converged"] >= 4 and i1["max_residual"] <= 1e-12


def test_eigsolve_reports_non_convergence(gpu_ctx):
    w = W.medium()
    sp = me.ED_momentum_subspaces(w.para, 9, momentum_restriction=True, momentum_list=[(0, 0)])[0][0]
    H = dev.create_gpu_linearmap(w.operator(), sp)
    vals, _, info = dev.krylov_map_solve(H, 6, krylovdim=12, maxiter=1, want_vectors=False)
    assert info["converged"] < 6 and info["iterations"] == 1


# ------------------------------------------------------------------ large sector, sampled
def test_large_sector_enumeration_and_sampled_matvec(gpu_ctx, oracle_mod):
    import torch
    w = W.large()
    op = w.operator()
    print("[large] enumerate", flush=True)
    b = dev.DeviceBasis.enumerate(gpu_ctx, _system(w.para, w.N_each), 0)
    assert b.D == 34770492 and b.rank_mode == 1
    basis = b.download()
    assert np.all(basis[1:] > basis[:-1])
    rng = np.random.default_rng(5)
    sample = rng.integers(0, b.D, 4096)
    pc = np.array([bin(int(v)).count("1") for v in basis[sample]])
    assert np.all(pc == 12)
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, w.para.bits)
    x = _rand(b.D, 12345)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    print("[large] matvec", flush=True)
    H = dev.ROCmLinearMap(dop, b)
    # the default for this sector: the LDS rank-table kernel (14 + 22 bit split) at 1024 rows
    assert H.kernel() == 1024 | dev.TUNE_NARROW
    H(yd, xd)
    y = yd.cpu().numpy()
    # bitwise the global-table kernel's y, and both kernels' y the same in dispatch order as in
    # XCD chunks (34k row blocks: 66 full groups of 512 blocks and a tail)
    y2 = torch.empty_like(xd)
    os.environ["EDM_XCD_CHUNK"] = "1"
    try:
        H(y2, xd)
        assert torch.equal(y2, yd)
    finally:
        del os.environ["EDM_XCD_CHUNK"]
    # the plain walk (the default gathers x into LDS slots while it walks on)
    os.environ["EDM_NARROW_PIPE"] = "0"
    try:
        H(y2, xd)
        assert torch.equal(y2, yd)
    finally:
        del os.environ["EDM_NARROW_PIPE"]
    os.environ["EDM_MATVEC"] = "u1"
    os.environ["EDM_MATVEC_LANES"] = "1"
    try:
        assert H.kernel() == 512
        H(y2, xd)
        assert torch.equal(y2, yd)
        os.environ["EDM_XCD_CHUNK_LDSU"] = "1"
        H(y2, xd)
        assert torch.equal(y2, yd)
    finally:
        del os.environ["EDM_MATVEC"]
        del os.environ["EDM_MATVEC_LANES"]
        os.environ.pop("EDM_XCD_CHUNK_LDSU", None)
    rows = np.concatenate([np.arange(64), sample, np.arange(b.D - 64, b.D)])
    full = _oracle_full(oracle_mod, op)
    print("[large] oracle rows", flush=True)
    yo = oracle_mod.matvec(basis, full, x, rows=rows)
    assert np.abs(y[rows] - yo).max() <= Y_TOL * np.abs(yo).max()


# ------------------------------------------------------------------ 48 orbitals (rank tables)
def test_literal_batched_sector_48_bits(gpu_ctx, oracle_mod):
    """The batched configuration as written: two conserved components on a 4x6 mesh each (48
    orbitals), N_up = N_down = 4, sector (0,0) (about 4.7 M states). Its basis gets rank
    tables (24 + 24 bits) and the two-level 14-bit split of the LDS rank-table kernel; H x is
    checked against the oracle on sampled rows and is bitwise the same with the global-table
    kernel, the plain walk, 512-row workgroups and the bucketed binary search (rank tables
    dropped)."""
    import torch
    para = W.literal_batched_para(1.3)
    op = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    sysd = _system(para, (4, 4))
    print("[48-bit] enumerate sector (0,0)", flush=True)
    b = dev.DeviceBasis.enumerate(gpu_ctx, sysd, 0)
    assert b.rank_mode == 1 and 4_000_000 < b.D < 5_500_000
    basis = b.download()
    assert np.all(basis[1:] > basis[:-1])
    rng = np.random.default_rng(48)
    sample = np.unique(np.concatenate([np.arange(32), rng.integers(0, b.D, 4096),
                                       np.arange(b.D - 32, b.D)]))
    lo = basis[sample] & np.uint64((1 << 24) - 1)
    hi = basis[sample] >> np.uint64(24)
    assert all(bin(int(v)).count("1") == 4 for v in lo)
    assert all(bin(int(v)).count("1") == 4 for v in hi)
    k = para.k_list
    for s in basis[sample[:256]]:
        bits = [i for i in range(48) if (int(s) >> i) & 1]
        ks = np.sum(k[:, [i % 24 for i in bits]], axis=1) % np.array([4, 6])
        assert tuple(ks) == (0, 0)
    dop = dev.DeviceOperator.from_mboperator(gpu_ctx, op, para.bits)
    x = _rand(b.D, 4848)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    print("[48-bit] matvec with rank tables", flush=True)
    H = dev.ROCmLinearMap(dop, b)
    # the default: the LDS rank-table kernel on the two-level split (14-bit low table in LDS,
    # the 34-bit high part as a 24-bit top table and 10-bit rows), bitwise the global-table
    # kernel and its own plain walk
    assert H.kernel() == 1024 | dev.TUNE_NARROW
    H(yd, xd)
    y = yd.cpu().numpy()
    y1 = torch.empty_like(xd)
    for env in ({"EDM_NARROW_PIPE": "0"}, {"EDM_MATVEC": "u1", "EDM_MATVEC_LANES": "1"},
                {"EDM_MATVEC": "u1n512"}):
        os.environ.update(env)
        try:
            if "EDM_MATVEC_LANES" in env:
                assert H.kernel() == 512
            H(y1, xd)
            assert torch.equal(y1, yd), env
        finally:
            for k in env:
                del os.environ[k]
    full = _oracle_full(oracle_mod, op)
    print("[48-bit] oracle rows", flush=True)
    yo = oracle_mod.matvec(basis, full, x, rows=sample)
    assert np.abs(y[sample] - yo).max() <= Y_TOL * np.abs(yo).max()
    b.drop_rank()
    assert b.rank_mode == 0
    print("[48-bit] matvec with the bucketed binary search", flush=True)
    y2 = torch.empty_like(xd)
    dev.ROCmLinearMap(dop, b)(y2, xd)
    assert torch.equal(y2, yd)
    # a small sector of the same 48-bit system (N_up = N_down = 1): tables of 2 x 2^24 words
    # would dwarf its few states, so it keeps the bucketed search, with the same result as
    # the oracle on every row
    sys1 = _system(para, (1, 1))
    b1 = dev.DeviceBasis.enumerate(gpu_ctx, sys1, 0)
    assert b1.rank_mode == 0 and 0 < b1.D < 1000
    basis1 = b1.download()
    x1 = _rand(b1.D, 11)
    y1 = dev.ROCmLinearMap(dop, b1)(x1)
    yo1 = oracle_mod.matvec(basis1, full, x1)
    assert np.abs(y1 - yo1).max() <= Y_TOL * np.abs(yo1).max()


def _random_conserving_op(rng, k, dims, ncomp, n2=240, n1=12):
    """A general (not Hermitian) operator of random complex two-body terms that conserve
    momentum (k_a + k_b = k_c + k_d on the mesh) and each component's particle number, plus
    number operators with complex amplitudes; orbital comp * nk + i has momentum k[:, i]."""
    nk = k.shape[1]
    dims = np.asarray(dims)
    where = {tuple(k[:, i] % dims): i for i in range(nk)}
    bits = ncomp * nk
    terms = []
    while len(terms) < n2:
        c1, c2 = rng.integers(ncomp, size=2)
        i1, i2, o1 = rng.integers(nk, size=3)
        o2 = where[tuple((k[:, i1] + k[:, i2] - k[:, o1]) % dims)]
        d, c = c1 * nk + i1, c2 * nk + i2
        a, b_ = c1 * nk + o1, c2 * nk + o2
        if a == b_ or c == d:
            continue
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a) + 1, int(b_) + 1,
                                int(c) + 1, int(d) + 1, bits=bits))
    for a in rng.choice(bits, n1, replace=False):
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a) + 1, int(a) + 1,
                                bits=bits))
    return me.MBOperator(terms, upper_hermitian=False)


@pytest.mark.parametrize("case", ["39 bits", "50 bits"])
def test_two_level_rank_split_edges(gpu_ctx, oracle_mod, monkeypatch, case):
    """The two-level split of the LDS rank-table kernel at both ends of its range: a 39-bit
    basis (3x13 mesh, N = 7: rows of 2 words under the 24-bit top table) and a 50-bit one (two
    conserved components on 5x5 meshes, N = 6 + 3: rows of 4,096 words). A random momentum-
    and component-conserving general operator; y of the one-lane global-table kernel is the
    oracle's to 1e-13 on sampled rows, and the LDS rank-table kernel (1024- and 512-row
    workgroups; pipelined, plain and forced-pipelined walk), the block kernel and the bucketed
    binary search give it bit for bit."""
    import torch
    if case == "39 bits":
        k, dims, ncomp, n_each = W.mesh(3, 13), (3, 13), 1, (7,)
    else:
        k, dims, ncomp, n_each = W.mesh(5, 5), (5, 5), 2, (6, 3)
    nbits = ncomp * k.shape[1]
    op = _random_conserving_op(np.random.default_rng(nbits), k, dims, ncomp)
    print(f"[split edges] {case}: enumerate", flush=True)
    b = dev.DeviceBasis.enumerate(gpu_ctx, dev.system_struct(k, dims, 1, ncomp, n_each), 0)
    assert b.rank_mode == 1 and b.D >= 1 << 18, (b.rank_mode, b.D)
    d = dev.DeviceOperator.from_mboperator(gpu_ctx, op, nbits)
    H = dev.ROCmLinearMap(d, b)
    x = _rand(b.D, 4000 + nbits)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    monkeypatch.setenv("EDM_MATVEC", "u1")
    monkeypatch.setenv("EDM_MATVEC_LANES", "1")
    assert H.kernel() & 0xFFFF == 512 and not H.kernel() & dev.TUNE_NARROW
    H(yd, xd)
    monkeypatch.delenv("EDM_MATVEC_LANES")
    y = yd.cpu().numpy()
    basis = b.download()
    rows = np.unique(np.concatenate([np.arange(64), np.arange(b.D - 64, b.D),
                                     np.random.default_rng(5).integers(0, b.D, 4000)]))
    print(f"[split edges] {case}: D={b.D}, oracle rows", flush=True)
    yo = oracle_mod.matvec(basis, _oracle_full(oracle_mod, op), x, rows=rows)
    assert np.count_nonzero(yo) > len(rows) // 2
    assert np.abs(y[rows] - yo).max() <= Y_TOL * np.abs(yo).max()
    y1 = torch.empty_like(xd)
    for v, blk in (("u1n", 1024), ("u1n512", 512)):
        monkeypatch.setenv("EDM_MATVEC", v)
        assert H.kernel() == blk | dev.TUNE_NARROW, (case, v, hex(H.kernel()))
        for pipe in (None, "0", "2"):
            if pipe is None:
                monkeypatch.delenv("EDM_NARROW_PIPE", raising=False)
            else:
                monkeypatch.setenv("EDM_NARROW_PIPE", pipe)
            H(y1, xd)
            assert torch.equal(y1, yd), (case, v, pipe)
        monkeypatch.delenv("EDM_NARROW_PIPE")
    monkeypatch.setenv("EDM_MATVEC", "u1")
    xs = [np.ascontiguousarray(x * (v + 1) + _rand(b.D, 77 + v)) for v in range(2)]
    yb = _gpu_matvec_block(gpu_ctx, d, b, xs)
    for v in range(2):
        np.testing.assert_array_equal(yb[v], _gpu_matvec(gpu_ctx, d, b, xs[v]),
                                      err_msg=f"{case} block col {v}")
    monkeypatch.delenv("EDM_MATVEC")
    b.drop_rank()
    assert b.rank_mode == 0
    print(f"[split edges] {case}: bucketed binary search", flush=True)
    dev.ROCmLinearMap(d, b)(y1, xd)
    assert torch.equal(y1, yd)


# ------------------------------------------------------------------ general eigensolver
def _random_general_op(rng, bits, n2=60, n1=20):
    terms = []
    for _ in range(n2):
        a, b_, c, d = rng.choice(bits, 4, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a), int(b_), int(c),
                                int(d), bits=bits))
    for _ in range(n1):
        a, c = rng.choice(bits, 2, replace=False) + 1
        terms.append(me.Scatter(complex(rng.normal(), rng.normal()), int(a), int(c), bits=bits))
    return me.MBOperator(terms, upper_hermitian=False)


def test_eigsolve_general_non_hermitian(gpu_ctx, oracle_mod):
    """ishermitian=False (Krylov-Schur on the device, as KrylovKit runs for a non-Hermitian
    map, linear_map.jl:179-191) on a general operator of 10 orbitals and 4 particles: the 6
    eigenvalues of smallest real part equal LAPACK's (numpy.linalg.eig on the oracle's dense
    matrix) to 1e-10, right eigenvectors have residual <= 1e-9; the same through
    EDsolve(method="map", ishermitian=False) and through the stored matrix
    (krylov_matrix_solve(..., ishermitian=False)); a Hermitian operator solved as a general
    one gives the Hermitian solver's (real) eigenvalues."""
    rng = np.random.default_rng(29)
    bits = 10
    op = _random_general_op(rng, bits)
    full = _oracle_full(oracle_mod, op)
    basis = np.array([s for s in range(1 << bits) if bin(s).count("1") == 4], dtype=np.uint64)
    Hd = oracle_mod.dense_matrix(basis, full)
    ev = np.linalg.eigvals(Hd)
    ev = ev[np.lexsort((ev.imag, ev.real))]
    scale = np.abs(ev).max()
    sp = me.HilbertSubspace(basis, bits=bits)
    lm = dev.create_gpu_linearmap(op, sp)
    vals, vecs, info = dev.krylov_map_solve(lm, 6, ishermitian=False, krylovdim=30, tol=1e-12)
    print(f"[general] {info}", flush=True)
    assert info["converged"] == 6 and vals.dtype == np.complex128
    np.testing.assert_allclose(vals, ev[:6], rtol=0, atol=1e-10 * scale)
    for lam, v in zip(vals, vecs):
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert np.linalg.norm(Hd @ v - lam * v) <= 1e-9 * scale
    E, V = me.EDsolve(sp, op, N=6, method="map", ishermitian=False, method_info=False)
    np.testing.assert_array_equal(E, vals)
    with pytest.raises(AssertionError):
        me.EDsolve(sp, op, N=6, method="sparse", ishermitian=False, method_info=False)
    Hm = dev.create_gpu_matrix(op, sp)
    vm, _, im = dev.krylov_matrix_solve(Hm, 6, ishermitian=False, krylovdim=30)
    Hm.close()
    assert im["converged"] == 6
    np.testing.assert_allclose(vm, ev[:6], rtol=0, atol=1e-10 * scale)
    # Hermitian operator through the general solver
    w = W.small()
    sps = me.ED_momentum_subspaces(w.para, 5)[0][0]
    Hh = dev.create_gpu_linearmap(w.operator(), sps)
    vh, _, _ = dev.krylov_map_solve(Hh, 8, want_vectors=False)
    vg, _, ig = dev.krylov_map_solve(Hh, 8, ishermitian=False, want_vectors=False)
    assert ig["converged"] == 8
    np.testing.assert_allclose(vg.real, vh, rtol=1e-10, atol=0)
    assert np.abs(vg.imag).max() <= 1e-10 * np.abs(vh).max()
