"""Batched and continuously batched solves (libedm's batch engine) through the C-ABI.

Bar: every problem's eigenvalues and eigenvectors are bitwise those of the same solve run
alone by edm_eigsolve (the batch engine runs the identical iteration with kernels of the same
per-problem arithmetic and reduction order), whatever shares the batch and whenever the
problem was admitted; against dense LAPACK on the oracle's matrix within 1e-10 relative.
"""
import numpy as np
import pytest

import momentumed as me
from momentumed import batch as B
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import EDM_ENOMEM, EDMError

pytestmark = pytest.mark.gpu


def _oracle_full(O, op):
    amp, out, inn = op.arrays()
    return O.full_list(O.TermList(amp, out, inn), op.upper_hermitian)


def _alone(H, nev, **kw):
    vals, vecs, info = dev.krylov_map_solve(H, nev, **kw)
    return vals, np.array(vecs), info


def _two_component_problems(ctx, ratios, fillings, sectors):
    """(map, label) for each (ratio, filling, sector): operators shared per ratio, bases per
    (filling, sector)."""
    out = []
    spaces = {}
    for f in fillings:
        subs, ks = me.ED_momentum_subspaces(W.batched_para(1.0), f)
        for s in sectors:
            spaces[(f, s)] = subs[s]
    for r in ratios:
        para = W.batched_para(r)
        op = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
        dop = dev.DeviceOperator.from_mboperator(ctx, op, 24)
        for f in fillings:
            for s in sectors:
                sp = spaces[(f, s)]
                out.append((dev.ROCmLinearMap(dop, dev.device_basis_of(sp, ctx), sp),
                            (r, f, s), op))
    return out


def test_batch_small_all_sectors_bitwise_equal_alone(gpu_ctx, oracle_mod):
    print("[batch] small 3x5 N=5, 15 sectors, nev 8", flush=True)
    w = W.small()
    op = w.operator()
    subspaces, _ = me.ED_momentum_subspaces(w.para, 5)
    maps = [dev.create_gpu_linearmap(op, sp) for sp in subspaces]
    with B.BatchSolver(gpu_ctx) as bs:
        ids = [bs.add(m, w.nev, want_vectors=True, seed=100 + i) for i, m in enumerate(maps)]
        res = bs.run()
    full = _oracle_full(oracle_mod, op)
    for i, (m, pid) in enumerate(zip(maps, ids)):
        r = res[pid]
        vals, vecs, info = _alone(m, w.nev, seed=100 + i)
        assert r.converged
        np.testing.assert_array_equal(r.values, vals)
        np.testing.assert_array_equal(np.array(r.vectors), vecs)
        assert r.info["matvecs"] == info["matvecs"]
        assert r.info["iterations"] == info["iterations"]
        ev = np.linalg.eigvalsh(oracle_mod.dense_matrix(subspaces[i].list, full))
        np.testing.assert_allclose(r.values, ev[: w.nev], rtol=1e-10, atol=1e-12)


def test_batch_two_component_uneven_sizes(gpu_ctx, oracle_mod):
    """Different operators (inter-component ratio) and bases (fillings, sectors) in one batch."""
    print("[batch] two-component 3x4, 3 ratios x 3 fillings x 2 sectors, nev 4", flush=True)
    probs = _two_component_problems(gpu_ctx, [0.6, 1.0, 1.9], [(2, 3), (3, 3), (4, 4)], [0, 7])
    with B.BatchSolver(gpu_ctx) as bs:
        ids = [bs.add(m, W.BATCH_NEV, want_vectors=True) for m, _, _ in probs]
        res = bs.run()
    checked = 0
    for (m, label, op), pid in zip(probs, ids):
        r = res[pid]
        vals, vecs, _ = _alone(m, W.BATCH_NEV)
        assert r.converged, label
        np.testing.assert_array_equal(r.values, vals)
        np.testing.assert_array_equal(np.array(r.vectors), vecs)
        if label[1] == (2, 3) and checked < 2:     # dense check on the smallest sectors
            ev = np.linalg.eigvalsh(oracle_mod.dense_matrix(m.space.list, _oracle_full(oracle_mod, op)))
            np.testing.assert_allclose(r.values, ev[: W.BATCH_NEV], rtol=1e-10, atol=1e-12)
            checked += 1
    assert checked == 2


@pytest.mark.parametrize("interleave", [None, "0"])
def test_batch_shared_walk_groups_bitwise_equal_alone(gpu_ctx, monkeypatch, interleave):
    """Nine parameter points on each of two bases: the batch walks each basis's rows once
    for up to four problems (groups of 4, 4 and 1 per basis); every problem's values,
    vectors and matvec count are still those of its lone solve. Also with every group
    gathering x from an interleaved copy of its vectors (by default only groups of 2^18 rows
    and more do)."""
    if interleave is not None:
        monkeypatch.setenv("EDM_BATCH_INTERLEAVE_MIN", interleave)
    ratios = [0.55, 0.7, 0.85, 1.1, 1.25, 1.4, 1.6, 1.8, 1.95]
    probs = _two_component_problems(gpu_ctx, ratios, [(3, 4)], [0, 5])
    with B.BatchSolver(gpu_ctx) as bs:
        ids = [bs.add(m, W.BATCH_NEV, want_vectors=True) for m, _, _ in probs]
        res = bs.run()
    for (m, label, _), pid in zip(probs, ids):
        r = res[pid]
        vals, vecs, info = _alone(m, W.BATCH_NEV)
        np.testing.assert_array_equal(r.values, vals)
        np.testing.assert_array_equal(np.array(r.vectors), vecs)
        assert r.info["matvecs"] == info["matvecs"], label


def test_batch_big_groups_bitwise_equal_alone(gpu_ctx, monkeypatch):
    """Groups of 2^18 rows and more walk their rows in a launch of their own (1,024-row
    workgroups, the 14-bit LDS rank split in either form): four ratios of the 48-orbital
    two-component sector (D = 4.7 M, the two-level split), then small problems forced through
    the same launch; every problem's values, vectors and matvec count are its lone solve's."""
    para0 = W.literal_batched_para(1.0)
    sysd = dev.system_struct(para0.k_list, para0.Gk, para0.Nc_mix, para0.Nc_conserve, (4, 4))
    b = dev.DeviceBasis.enumerate(gpu_ctx, sysd, 0)
    maps = []
    for r in (0.6, 0.9, 1.3, 1.7):
        pr = W.literal_batched_para(r)
        op = me.MBOperator(me.ED_scatterlist_twobody(pr), upper_hermitian=True)
        maps.append(dev.ROCmLinearMap(dev.DeviceOperator.from_mboperator(gpu_ctx, op, pr.bits), b))
    kw = dict(krylovdim=16, tol=1e-6)
    print(f"[big-group] 48-orbital sector D={b.D}: batch of 4", flush=True)
    with B.BatchSolver(gpu_ctx) as bs:
        ids = [bs.add(m, 2, **kw) for m in maps]
        res = bs.run()
    for m, pid in zip(maps, ids):
        vals, _, info = dev.krylov_map_solve(m, 2, want_vectors=False, **kw)
        np.testing.assert_array_equal(res[pid].values, vals)
        assert res[pid].info["matvecs"] == info["matvecs"]
    monkeypatch.setenv("EDM_BATCH_BIG_GROUP_ROWS", "0")
    probs = _two_component_problems(gpu_ctx, [0.55, 0.7, 0.85, 1.1, 1.25], [(3, 4)], [0, 5])
    print("[big-group] small problems through the big-group launch", flush=True)
    with B.BatchSolver(gpu_ctx) as bs:
        ids = [bs.add(m, W.BATCH_NEV, want_vectors=True) for m, _, _ in probs]
        res = bs.run()
    for (m, label, _), pid in zip(probs, ids):
        vals, vecs, info = _alone(m, W.BATCH_NEV)
        np.testing.assert_array_equal(res[pid].values, vals)
        np.testing.assert_array_equal(np.array(res[pid].vectors), vecs)
        assert res[pid].info["matvecs"] == info["matvecs"], label


def test_batch_continuous_admission_and_removal(gpu_ctx):
    """Problems admitted while others are mid-solve, one removed mid-run: every collected
    result equals its solve alone."""
    print("[batch] continuous admission", flush=True)
    probs = _two_component_problems(gpu_ctx, [0.8, 1.5], [(2, 4), (3, 4)], [1, 5])
    with B.BatchSolver(gpu_ctx) as bs:
        ids = {}
        for i in range(3):
            ids[i] = bs.add(probs[i][0], W.BATCH_NEV, want_vectors=True, seed=i)
        bs.step(7)
        for i in range(3, 6):
            ids[i] = bs.add(probs[i][0], W.BATCH_NEV, want_vectors=True, seed=i)
        bs.step(3)
        bs.remove(ids.pop(4))
        for i in range(6, len(probs)):
            ids[i] = bs.add(probs[i][0], W.BATCH_NEV, want_vectors=True, seed=i)
        res = bs.run(max_steps_per_call=5)
        assert bs.count() == (0, 0)
    assert sorted(res) == sorted(ids.values())
    for i, pid in ids.items():
        vals, vecs, _ = _alone(probs[i][0], W.BATCH_NEV, seed=i)
        np.testing.assert_array_equal(res[pid].values, vals)
        np.testing.assert_array_equal(np.array(res[pid].vectors), vecs)


def test_batch_not_converged_and_vec0_match_alone(gpu_ctx):
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][3]
    H = dev.create_gpu_linearmap(op, sp)
    rng = np.random.default_rng(5)
    v0 = rng.normal(size=sp.length) + 1j * rng.normal(size=sp.length)
    with B.BatchSolver(gpu_ctx) as bs:
        a = bs.add(H, 6, krylovdim=10, maxiter=1, want_vectors=True)
        b = bs.add(H, 3, vec0=v0, want_vectors=False)
        res = bs.run()
    assert not res[a].converged and res[a].info["iterations"] == 1
    vals, vecs, info = _alone(H, 6, krylovdim=10, maxiter=1)
    assert info["converged"] < 6
    np.testing.assert_array_equal(res[a].values, vals)
    np.testing.assert_array_equal(np.array(res[a].vectors), vecs)
    vals, _, _ = dev.krylov_map_solve(H, 3, vec0=v0, want_vectors=False)
    np.testing.assert_array_equal(res[b].values, vals)
    assert res[b].vectors is None


def test_breakdown_mid_cycle_lone_and_batched(gpu_ctx):
    """A start vector inside a three-dimensional invariant subspace (three basis states of a
    diagonal one-body operator with distinct diagonal entries): the Krylov space runs out at
    the third step of the first cycle, exactly (every vector stays on those three states), so
    the lone solve discards the step it had already queued behind that one and continues from
    a random direction; the lowest four eigenvalues are then the four smallest diagonal
    entries, and the batched solve (whose host step handles the same breakdown) is bitwise
    the lone one."""
    print("[batch] breakdown at step 3: diagonal 3x5 N=5 operator, vec0 on three states",
          flush=True)
    eps = [0.37 * np.sqrt(k + 2.0) + 0.011 * k * k for k in range(15)]
    para = me.EDPara(k_list=W.mesh(3, 5), Gk=(3, 5),
                     H_one=lambda k, cf, ci: eps[k - 1] if cf == ci else 0.0,
                     one_momentum_coordinate=False)
    op = me.MBOperator(me.ED_scatterlist_onebody(para), upper_hermitian=True)
    sp = me.ED_momentum_subspaces(para, 5)[0][0]
    states = np.asarray(sp.list, dtype=np.uint64)
    occ = (states[:, None] >> np.arange(15, dtype=np.uint64)) & np.uint64(1)
    diag = occ.astype(float) @ np.asarray(eps)
    order = np.argsort(diag)
    assert sp.length > 100 and np.min(np.diff(diag[order])) > 1e-9
    v0 = np.zeros(sp.length, complex)
    v0[order[-3:]] = 1.0            # the three largest entries: the answer needs the restart
    H = dev.create_gpu_linearmap(op, sp)
    vals, vecs, info = _alone(H, 4, vec0=v0)
    assert info["converged"] == 4
    np.testing.assert_allclose(vals, diag[order][:4], rtol=1e-10, atol=1e-12)
    for i in range(4):
        assert abs(abs(vecs[i][order[i]]) - 1.0) < 1e-9
    with B.BatchSolver(gpu_ctx) as bs:
        pid = bs.add(H, 4, vec0=v0, want_vectors=True)
        res = bs.run()
    np.testing.assert_array_equal(res[pid].values, vals)
    np.testing.assert_array_equal(np.array(res[pid].vectors), vecs)
    assert res[pid].info["matvecs"] == info["matvecs"]


def _map_on(ctx, op, sp):
    system, sector = sp.sector_info
    basis = dev.DeviceBasis.upload(ctx, sp.list, sp.bits, system, sector)
    return dev.ROCmLinearMap(dev.DeviceOperator.from_mboperator(ctx, op, sp.bits), basis, sp)


def _drop(H):
    H.op.close()
    H.basis.close()


def test_batch_budget_admission(gpu_ctx):
    """A context with an HBM budget refuses a problem that does not fit, and admits it once
    a finished problem has been collected."""
    print("[batch] HBM budget admission", flush=True)
    w = W.small()
    op = w.operator()
    sp = me.ED_momentum_subspaces(w.para, 5)[0][0]
    ctx = dev.Context(0)
    used0, _ = ctx.memory()
    H = _map_on(ctx, op, sp)
    used1, _ = ctx.memory()
    _drop(H)
    ctx.close()
    need = B.problem_bytes(sp.length, 4, 30)
    # room for the map, one problem and the batch's small staging buffers, not two problems
    budget = used1 + need + need // 2
    ctx = dev.Context(0, hbm_budget_bytes=budget)
    H = _map_on(ctx, op, sp)
    try:
        with B.BatchSolver(ctx) as bs:
            a = bs.add(H, 4)
            with pytest.raises(EDMError) as e:
                bs.add(H, 4)
            assert e.value.code == EDM_ENOMEM
            while bs.step(50)[0]:
                pass
            r = bs.result(a)
            b = bs.add(H, 4)
            res = bs.run()
        np.testing.assert_array_equal(res[b].values, r.values)
        vals, _, _ = dev.krylov_map_solve(H, 4, want_vectors=False)
        np.testing.assert_array_equal(r.values, vals)
    finally:
        _drop(H)
        ctx.close()


def test_edsolve_batch_api(gpu_ctx):
    w = W.small()
    op = w.operator()
    subspaces, _ = me.ED_momentum_subspaces(w.para, 5)
    pairs = [(sp, op) for sp in subspaces[:4]]
    out = me.EDsolve_batch(pairs, N=4)
    for (sp, _), (E, V) in zip(pairs, out):
        E1, V1 = me.EDsolve(sp, op, N=4, method="map", device="gpu", method_info=False)
        np.testing.assert_array_equal(E, E1)
        for a, b in zip(V, V1):
            np.testing.assert_array_equal(a.vec, b.vec)


def test_served_results_equal_alone(gpu_ctx):
    """Continuous batching: requests arriving while others run; each served result is bitwise
    the request solved alone."""
    from momentumed import serve as S
    print("[serve] 0.6 s Poisson stream of 3x5/4x5/4x6 requests at 150/s", flush=True)
    reqs = S.served_stream(rate=150.0, duration=0.6, seed=11, meshes=((3, 5), (4, 5), (4, 6)))
    srv = S.SolveServer(hbm_budget_bytes=8 << 30, max_queue=1000)
    try:
        rep = srv.run(reqs, drain_s=120.0)
    finally:
        srv.close()
    summ = rep.summary()
    print(summ, flush=True)
    assert rep.completed == len(reqs) and rep.rejected == 0 and rep.unfinished == 0
    assert rep.steps > 0
    for r in reqs:
        np.testing.assert_array_equal(rep.results[r.rid], S.solve_alone(r, gpu_ctx))


@pytest.mark.timeout(900)
def test_served_stream_with_large_requests(gpu_ctx, oracle_mod):
    """A served stream holding 5x6 requests and a 6x6@1/3 request (34.8 M states) among small
    ones, on the server's default scheduling: every request completes (the large one is
    not starved by the small ones), each result is bitwise the request solved alone, and
    the values are exact eigenvalues: small sectors against the oracle's matrix (dense or
    ARPACK), the 5x6 and 6x6 ones through oracle-sampled residuals of their eigenvectors."""
    from fractions import Fraction

    import scipy.sparse as sps2
    import scipy.sparse.linalg as spl
    from momentumed import serve as S
    print("[serve] mixed stream with 5x6 and 6x6 requests", flush=True)
    small = S.served_stream(rate=60.0, duration=0.25, seed=21,
                            meshes=((3, 5), (4, 5), (4, 6)))
    n0 = len(small)
    big = [S.Request(n0, 0.01, (6, 6), Fraction(1, 3), 12, (1, 2), 2.5),
           S.Request(n0 + 1, 0.03, (5, 6), Fraction(1, 3), 10, (2, 3), 1.7),
           S.Request(n0 + 2, 0.06, (5, 6), Fraction(2, 5), 12, (4, 1), 3.1)]
    reqs = small + big
    srv = S.SolveServer(hbm_budget_bytes=200 * 10 ** 9, max_queue=1000)
    try:
        rep = srv.run(reqs, drain_s=600.0)
    finally:
        srv.close()
    summ = rep.summary()
    print({k: summ[k] for k in ("completed", "unfinished", "elapsed_s", "steps")},
          summ["by_class"], flush=True)
    assert rep.completed == len(reqs) and rep.rejected == 0 and rep.unfinished == 0
    for r in reqs:
        want_vec = r.mesh in ((5, 6), (6, 6))
        print(f"[serve] alone: request {r.rid} {r.label} sector {r.sector}", flush=True)
        got = S.solve_alone(r, gpu_ctx, want_vectors=want_vec)
        vals = got[0] if want_vec else got
        np.testing.assert_array_equal(rep.results[r.rid], vals)
        amp, out, inn = S.request_operator(r).arrays()
        full = oracle_mod.full_list(oracle_mod.TermList(amp, out, inn), True)
        if want_vec:
            _, vecs, states, info = got
            assert info["converged"] == r.nev
            rng = np.random.default_rng(r.rid)
            rows = rng.integers(0, len(states), 2048)
            for i, v in enumerate(vecs):
                yo = oracle_mod.matvec(states, full, v, rows=rows)
                est = np.linalg.norm(yo - vals[i] * v[rows]) * np.sqrt(len(states) / len(rows))
                assert est <= 1e-9, (r.label, i, est)
            continue
        lists, ks = oracle_mod.enumerate_sectors(S.mesh(*r.mesh), r.mesh, 1, 1, [r.N])
        basis = lists[ks.index(tuple(r.sector))]
        if len(basis) <= 1500:
            ev = np.linalg.eigvalsh(oracle_mod.dense_matrix(basis, full))[: r.nev + 4]
        else:
            rp, cols, terms, negs = oracle_mod.connectivity(basis, full)
            Hs = sps2.csr_matrix((full.amp[terms] * np.where(negs, -1.0, 1.0), cols, rp),
                                 shape=(len(basis), len(basis)))
            ev = np.sort(spl.eigsh(Hs, k=r.nev + 4, which="SA", tol=1e-14, ncv=60)[0])
        # single-vector Lanczos may return one copy of a degenerate level: every value must
        # be an exact eigenvalue, and the lowest must be the ground state
        scale = np.abs(ev).max()
        for e in vals:
            assert np.abs(ev - e).min() <= 1e-10 * scale, (r.label, e, ev)
        assert abs(vals[0] - ev[0]) <= 1e-10 * scale
