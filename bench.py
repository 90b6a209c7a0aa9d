#!/usr/bin/env python3
"""Matrix-free Hamiltonian matvec throughput on one MI355X.

A step is one y = H x over a whole momentum sector, x resident in HBM. Default workload: the
6x6 mesh, N=12 sector (0,0) (34,770,492 states, 556 MB per complex vector). Prints one JSON
line (metric, value in matvecs/s, roofline of the matvec kernel from HIP events on the
library's stream, the CPU baseline on a bounded row sample, setup times). Progress goes to
stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|medium|small]
                    [--no-solve] [--no-cpu-baseline] [--cpu-seconds S] [--no-batched]
                    [--no-served] [--batch-ratios R] [--serve-seconds T]

Beside the matvec line it reports, in the same JSON object, the block matvec (4 vectors per
launch), the sector's lowest-nev eigensolve (time, matvec count, residual), the batched sweep
(3,840 two-component sector solves, batched vs a plain loop of single solves) and the served
stream (Poisson requests of mixed sizes, continuous batching vs one solve at a time, p50/p99
latency).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "momentumed.jl_amd"))

METRIC = "Hamiltonian matvecs/s & HBM GB/s; batched sector-solves/s + p99 latency vs plain"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _kernel_name(rank_mode: int, block: int = 512) -> str:
    """Name of the matvec kernel libedm.so launches for the bucketed path, from the code
    edm_matvec_kernel reports (rows per workgroup | 0x10000 for the LDS-amplitude kernel |
    0x20000 for the LDS rank-table kernel | lanes per row << 20 for the split-row kernel),
    EDM_MATVEC and EDM_NARROW_PIPE (the LDS rank-table kernel's pipelined walk, on unless it
    would cost occupancy; it is on for the workloads measured here)."""
    rank = "RankLin" if rank_mode == 1 else "RankSearch"
    v = os.environ.get("EDM_MATVEC", "u1")
    if block & 0x20000:
        pipe = os.environ.get("EDM_NARROW_PIPE", "1") != "0"
        return f"k_matvec_narrow<{block & 0xFFFF}, {str(pipe).lower()}>"
    if v == "u1" and (block >> 20) > 1:
        return f"k_matvec_split<{rank}, {(block >> 20) & 0xFF}>"
    if v == "u1a" or (v == "u1" and block & 0x10000):
        return f"k_matvec_ldsa<{rank}, {block & 0xFFFF if block & 0x10000 else 1024}>"
    if v == "basic" or block == 0:
        return f"k_matvec_fast<{rank}>"
    if v in ("hi", "lo"):
        return f"k_matvec_lds<{rank}, {str(v == 'hi').lower()}>"
    u = {"u2": 2, "u4": 4}.get(v, 1)
    return f"k_matvec_ldsu<{rank}, true, {u}, {block & 0xFFFF if u == 1 else 512}>"


def cpu_share():
    """Host threads the CPU baseline uses, and what it knows of the host: the threads this
    process may run on (sched_getaffinity), capped by OMP_NUM_THREADS when that is set (a
    GPU box of this pool gives one GPU's job a share of 16), nproc and the CPU model."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, {"threads": threads, "affinity_cpus": aff, "nproc": os.cpu_count(),
                     "omp_num_threads": omp, "cpu_model": model}


def cpu_baseline(w, basis, full_terms, x_host, y_dev_host, seconds, nthreads):
    """The oracle's row-gather matvec (restatement of linear_map.jl:100-118: every term tried
    on every row, binary search for partners) on a random row sample, scaled to matvecs/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    D = len(basis)
    rng = np.random.default_rng(2024)
    n = min(D, 512)
    rows = rng.choice(D, n, replace=False)
    t0 = time.perf_counter()
    yo = oracle.matvec(basis, full_terms, x_host, rows=rows, nthreads=nthreads)
    dt = time.perf_counter() - t0
    per_row = dt / n
    n2 = int(min(D, max(n, seconds / max(per_row, 1e-9))))
    if n2 > n:
        rows = rng.choice(D, n2, replace=False) if n2 < D else np.arange(D)
        t0 = time.perf_counter()
        yo = oracle.matvec(basis, full_terms, x_host, rows=rows, nthreads=nthreads)
        dt = time.perf_counter() - t0
        n = n2
    rows_per_s = n / dt
    err = float(np.abs(y_dev_host[rows] - yo).max() / max(np.abs(yo).max(), 1e-300))
    return {
        "value": rows_per_s / D, "unit": "matvec/s", "cores": nthreads, "kind": "port",
        "sample": f"{n} random rows of the {D}-state sector ({dt:.1f} s, OpenMP row-parallel "
                  f"restatement of linear_map.jl:100-118, all full-list terms per row, "
                  f"binary search), scaled to whole matvecs",
        "rows_per_s": rows_per_s,
    }, err


def block_section(ctx, dop, basis, x, y, stream, args, nnz_off, S, ms_single, narrow=False):
    """y_v = H x_v for 4 vectors per edm_matvec_block call (the block-Lanczos matvec: one
    candidate walk and one partner lookup serve all 4 vectors; x interleaved so a partner's 4
    entries share a cache line). Timed with HIP events on the library stream like the single
    matvec; algorithmic bytes D*(8 + 4*32) + 4*16*nnz_off + 40*S per call."""
    import torch
    from momentumed._lib import check
    nv, D = 4, basis.D
    X = [x] + [torch.empty_like(x) for _ in range(nv - 1)]
    for v in range(1, nv):
        check(ctx.L.edm_fill_random(ctx.h, X[v].data_ptr(), D, 12345 + v, 1), ctx.h)
    Y = [y] + [torch.empty_like(x) for _ in range(nv - 1)]
    work = torch.empty(nv * D, dtype=torch.complex128, device=x.device)
    ctx.synchronize()
    torch.cuda.synchronize()
    xp = (ctypes.c_void_p * nv)(*[t.data_ptr() for t in X])
    yp = (ctypes.c_void_p * nv)(*[t.data_ptr() for t in Y])

    def call():
        check(ctx.L.edm_matvec_block(ctx.h, dop.h, basis.h, nv, xp, yp, work.data_ptr()), ctx.h)
    for _ in range(max(1, args.warmup)):
        call()
    ctx.synchronize()
    steps = max(3, args.steps // 2)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        call()
    ev1.record(stream)
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    B = D * (8 + nv * 32) + nv * 16 * nnz_off + 40 * S
    ach = B / (ms * 1e-3) / 1e9
    log(f"[bench] block matvec x{nv}: {ms:.3f} ms per call, {ms / nv:.3f} ms per vector")
    del X[1:], Y[1:], work
    torch.cuda.empty_cache()
    return {"vectors_per_call": nv, "calls": steps, "ms_per_call": round(ms, 4),
            "matvecs_per_s": round(nv * 1e3 / ms, 3),
            "speedup_per_vector_vs_single": round(ms_single * nv / ms, 3),
            "kernel": ("k_matvec_block_narrow<4, 1024>" if narrow else
                       f"k_matvec_block<{'RankLin' if basis.rank_mode else 'RankSearch'}, 4, 512>")
                      + " + k_interleave<4>",
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                         "bytes_model": "D*(8+4*32) + 4*16*nnz_off + 40*S_full",
                         "algorithmic_bytes": B}}


def config_section(name, args):
    """The small or medium configuration solved on the GPU and on the host.

    small: all 15 sectors of the 3x5 mesh, N = 5, lowest 8 (krylovdim 30, tol 1e-12);
    medium: sector (0,0) of the 3x9 mesh, N = 9, lowest 6, krylovdim 40, tol 1e-12.
    GPU: krylov_map_solve per sector (operator and bases resident), sector-solves/s, and the
    matvec of the first sector timed with HIP events (GB/s of the bytes model). CPU: the
    oracle's restatement of the reference's map (linear_map.jl:100-118, OpenMP over rows)
    inside a thick-restart Lanczos (oracle/lanczos.py) on the host's CPU share; small: every
    sector solved; medium: one Krylov cycle (krylovdim steps) timed and scaled to the GPU
    solve's matvec count, labelled extrapolated."""
    import numpy as np
    import torch
    import momentumed as me
    from momentumed import device as dev
    from momentumed import workloads as W
    from momentumed._lib import check
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lanczos
    import oracle
    w = W.get(name)
    op = w.operator()
    if name == "small":
        spaces, ks = me.ED_momentum_subspaces(w.para, w.N_each)
    else:
        spaces, ks = me.ED_momentum_subspaces(w.para, w.N_each, momentum_restriction=True,
                                              momentum_list=[w.sector])
    ctx = dev.get_context(0)
    maps = [dev.create_gpu_linearmap(op, sp) for sp in spaces]
    kw = dict(krylovdim=w.krylovdim, tol=w.tol, want_vectors=False)
    dev.krylov_map_solve(maps[0], w.nev, **kw)   # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    gpu_vals, mvs = [], 0
    for H in maps:
        v, _, info = dev.krylov_map_solve(H, w.nev, **kw)
        gpu_vals.append(v)
        mvs += info["matvecs"]
    t_gpu = time.perf_counter() - t0
    H = maps[0]
    D = H.basis.D
    nnz, nnz_off = ctypes.c_int64(), ctypes.c_int64()
    check(ctx.L.edm_connectivity_totals(ctx.h, H.op.h, H.basis.h, ctypes.byref(nnz),
                                        ctypes.byref(nnz_off)), ctx.h)
    x = torch.empty(D, dtype=torch.complex128, device="cuda:0")
    y = torch.empty_like(x)
    check(ctx.L.edm_fill_random(ctx.h, x.data_ptr(), D, 12345, 1), ctx.h)
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream, device="cuda:0")
    reps = 200 if name == "small" else 50
    for _ in range(5):
        H.matvec_device(y.data_ptr(), x.data_ptr())
    ctx.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        H.matvec_device(y.data_ptr(), x.data_ptr())
    ev1.record(stream)
    ev1.synchronize()
    ms_mv = ev0.elapsed_time(ev1) / reps
    B = D * 40 + nnz_off.value * 16 + H.op.S_full * 40
    out = {"workload": w.name, "sectors": len(spaces), "D": [sp.length for sp in spaces],
           "nev": w.nev, "krylovdim": w.krylovdim, "tol": w.tol,
           "gpu_seconds": round(t_gpu, 4), "gpu_matvecs": mvs,
           "sector_solves_per_s": round(len(spaces) / t_gpu, 2),
           "matvec": {"sector": list(ks[0]), "ms": round(ms_mv, 5),
                      "algorithmic_bytes": B, "GB_per_s": round(B / (ms_mv * 1e-3) / 1e9, 1),
                      "nnz_offdiag": nnz_off.value,
                      "bytes_model": "D*40 + nnz_off*16 + S_full*40"}}
    log(f"[{name}] GPU: {len(spaces)} solves in {t_gpu:.3f} s, matvec {ms_mv:.4f} ms")
    if args.no_cpu_baseline:
        return out
    threads, host = cpu_share()
    amp, o, i = op.arrays()
    full = oracle.full_list(oracle.TermList(amp, o, i), True)
    cpu = {"kind": "port", "cores": threads, "host": host,
           "algorithm": "oracle restatement of linear_map.jl:100-118 (every full-list term "
                        "tried on every row, binary search) inside oracle/lanczos.py"}
    if name == "small":
        t0 = time.perf_counter()
        err = 0.0
        for sp, gv in zip(spaces, gpu_vals):
            b = np.ascontiguousarray(sp.list)
            vals, _, _ = lanczos.eigsolve(
                lambda v, b=b: oracle.matvec(b, full, v, nthreads=threads), len(b), w.nev,
                krylovdim=w.krylovdim, tol=w.tol)
            err = max(err, float(np.abs(vals - gv).max() / np.abs(gv).max()))
        t_cpu = time.perf_counter() - t0
        cpu.update({"seconds": round(t_cpu, 3), "sample": f"all {len(spaces)} sectors solved",
                    "sector_solves_per_s": round(len(spaces) / t_cpu, 3),
                    "eigenvalues_max_rel_diff_vs_gpu": err})
    else:
        b = np.ascontiguousarray(spaces[0].list)
        t0 = time.perf_counter()
        _, _, info = lanczos.eigsolve(lambda v: oracle.matvec(b, full, v, nthreads=threads),
                                      len(b), w.nev, krylovdim=w.krylovdim, tol=w.tol,
                                      maxiter=1)
        t_cyc = time.perf_counter() - t0
        per_step = t_cyc / info["matvecs"]
        t_cpu = per_step * mvs
        cpu.update({"seconds": round(t_cpu, 2), "extrapolated": True,
                    "sample": f"one {info['matvecs']}-step Lanczos cycle timed "
                              f"({t_cyc:.1f} s), scaled by the GPU solve's {mvs} matvecs",
                    "sector_solves_per_s": round(1.0 / t_cpu, 5)})
    cpu["gpu_speedup"] = round(t_cpu / t_gpu, 1)
    log(f"[{name}] CPU ({threads} threads): {cpu['seconds']} s")
    out["cpu"] = cpu
    return out


def batched_section(args):
    """Sector-solves/s of the batched sweep against the plain loop of single solves.

    Problems: every (ratio, filling, sector) of the two-component 3x4 system (64 ratios x 5
    fillings x 12 sectors), lowest 4 eigenvalues, krylovdim 30, tol 1e-12. The plain loop
    (krylov_map_solve per problem, operators and bases already on the device) is timed on the
    first ratio's 60 problems, which span every filling and sector; the batch solves all
    problems. The sample's eigenvalues must be bitwise equal in both."""
    import numpy as np
    import momentumed as me
    from momentumed import batch as B
    from momentumed import device as dev
    from momentumed import workloads as W
    ctx = dev.get_context(0)
    t0 = time.perf_counter()
    spaces = []
    for f in W.BATCH_FILLINGS:
        subs, ks = me.ED_momentum_subspaces(W.batched_para(1.0), f)
        for k, sp in zip(ks, subs):
            spaces.append((f, k, sp))
    bases = [dev.device_basis_of(sp, ctx) for _, _, sp in spaces]
    t_enum = time.perf_counter() - t0
    t0 = time.perf_counter()
    ratios = W.batched_ratios(args.batch_ratios)
    ops = []
    for r in ratios:
        op = me.MBOperator(me.ED_scatterlist_twobody(W.batched_para(r)), upper_hermitian=True)
        ops.append(dev.DeviceOperator.from_mboperator(ctx, op, 24))
    t_ops = time.perf_counter() - t0
    maps = [dev.ROCmLinearMap(o, b) for o in ops for b in bases]
    dims = [b.D for b in bases]
    nev, kd, tol = W.BATCH_NEV, 30, 1e-12
    log(f"[batched] {len(maps)} problems (D {min(dims)}..{max(dims)}), bases {t_enum:.2f} s, "
        f"operators {t_ops:.2f} s")
    nplain = min(len(bases), len(maps))
    # one untimed solve first, so the plain loop is timed warm (kernels loaded, the solver's
    # workspace and reduction buffers allocated) as the long batch run effectively is
    dev.krylov_map_solve(maps[0], nev, krylovdim=kd, tol=tol, want_vectors=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    plain_vals, plain_mv, plain_conv = [], 0, 0
    for m in maps[:nplain]:
        v, _, info = dev.krylov_map_solve(m, nev, krylovdim=kd, tol=tol, want_vectors=False)
        plain_vals.append(v)
        plain_mv += info["matvecs"]
        plain_conv += int(info["converged"] >= nev)
    t_plain = time.perf_counter() - t0
    log(f"[batched] plain loop: {nplain} solves in {t_plain:.2f} s ({nplain / t_plain:.1f}/s)")
    used0, _ = ctx.memory()
    peak = used0
    t0 = time.perf_counter()
    with B.BatchSolver(ctx) as bs:
        ids = [bs.add(m, nev, krylovdim=kd, tol=tol) for m in maps]
        t_add = time.perf_counter() - t0
        peak = max(peak, ctx.memory()[0])
        res = bs.run()
    t_batch = time.perf_counter() - t0
    nconv = sum(1 for r in res.values() if r.converged)
    mv = sum(r.info["matvecs"] for r in res.values())
    same = all(np.array_equal(res[ids[i]].values, plain_vals[i]) for i in range(nplain))
    log(f"[batched] batch: {len(maps)} solves in {t_batch:.2f} s ({len(maps) / t_batch:.1f}/s), "
        f"{nconv} converged, sample bitwise equal: {same}")
    sps_b, sps_p = len(maps) / t_batch, nplain / t_plain
    return {
        "workload": f"two-component 3x4 mesh, fillings {list(W.BATCH_FILLINGS)}, all 12 sectors, "
                    f"{len(ratios)} inter-component ratios ~U(0.5,2) seed 99",
        "problems": len(maps), "D_min": min(dims), "D_max": max(dims), "nev": nev,
        "krylovdim": kd, "tol": tol, "converged": nconv,
        "sector_solves_per_s": round(sps_b, 2), "wall_s": round(t_batch, 3),
        "converged_sector_solves_per_s": round(nconv / t_batch, 2),
        "plain_converged": plain_conv,
        "add_s": round(t_add, 3), "matvecs": mv,
        "plain_sector_solves_per_s": round(sps_p, 2), "plain_sample": nplain,
        "plain_sample_s": round(t_plain, 3), "plain_matvecs": plain_mv,
        "speedup_vs_plain": round(sps_b / sps_p, 2),
        "sample_bitwise_equal": bool(same),
        "workspace_bytes": int(peak - used0), "setup_s": round(t_enum + t_ops, 3),
    }


def literal_section(args):
    """One (sector, ratio) point of the batched configuration as written: two conserved
    components on a 4x6 mesh each (48 orbitals), N_up = N_down = 4, sector (0,0), ratio 1.3
    (D = 4,705,212): basis enumeration, matvec time, and the lowest 4 by the single-vector
    iteration (krylovdim 30) and by block Lanczos with block size 4 (krylovdim 60), each
    timed alone, with the sweep's time extrapolated from them (24 sectors x 64 ratios); and
    the batched form at this size: 8 ratios of the sector in one batch against single solves."""
    import torch
    import momentumed as me
    from momentumed import device as dev
    from momentumed import workloads as W
    para = W.literal_batched_para(1.3)
    op = me.MBOperator(me.ED_scatterlist_twobody(para), upper_hermitian=True)
    ctx = dev.get_context(0)
    sysd = dev.system_struct(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, (4, 4))
    t0 = time.perf_counter()
    b = dev.DeviceBasis.enumerate(ctx, sysd, 0)
    t_enum = time.perf_counter() - t0
    dop = dev.DeviceOperator.from_mboperator(ctx, op, para.bits)
    lm = dev.ROCmLinearMap(dop, b)
    x = torch.empty(b.D, dtype=torch.complex128, device="cuda:0")
    y = torch.empty_like(x)
    ctx.synchronize()
    from momentumed._lib import check
    check(ctx.L.edm_fill_random(ctx.h, x.data_ptr(), b.D, 12345, 1), ctx.h)
    for _ in range(3):
        lm.matvec_device(y.data_ptr(), x.data_ptr())
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream, device="cuda:0")
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(10):
        lm.matvec_device(y.data_ptr(), x.data_ptr())
    ev1.record(stream)
    ev1.synchronize()
    ms_mv = ev0.elapsed_time(ev1) / 10
    solves = {}
    for bs, kd in ((1, 30), (4, 60)):
        t0 = time.perf_counter()
        v, _, info = dev.krylov_map_solve(lm, 4, krylovdim=kd, tol=1e-12, blocksize=bs,
                                          want_vectors=False)
        ts = time.perf_counter() - t0
        solves[f"block{bs}"] = {"krylovdim": kd, "seconds": round(ts, 3),
                                "matvecs": info["matvecs"], "converged": info["converged"],
                                "eigenvalues": [float(e) for e in v]}
        log(f"[literal] block {bs}: {ts:.2f} s, {info['matvecs']} matvecs, values {v}")
    log(f"[literal] D={b.D} enumerated in {t_enum:.2f} s, matvec {ms_mv:.3f} ms")
    # the sweep's batched form at this size: the first 8 of the sweep's 64 ratios on the same
    # sector in one batch (two shared-walk groups of four) against single solves of the first 2
    import numpy as np
    from momentumed import batch as B
    ratios = W.batched_ratios(64)[:8]
    maps = []
    for r in ratios:
        pr = W.literal_batched_para(float(r))
        o = me.MBOperator(me.ED_scatterlist_twobody(pr), upper_hermitian=True)
        maps.append(dev.ROCmLinearMap(dev.DeviceOperator.from_mboperator(ctx, o, pr.bits), b))
    ctx.synchronize()
    t0 = time.perf_counter()
    plain = [dev.krylov_map_solve(m, 4, krylovdim=30, tol=1e-12, want_vectors=False)[0]
             for m in maps[:2]]
    t_plain = (time.perf_counter() - t0) / 2
    t0 = time.perf_counter()
    with B.BatchSolver(ctx) as bs:
        ids = [bs.add(m, 4, krylovdim=30, tol=1e-12) for m in maps]
        res = bs.run()
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(res[ids[i]].values, plain[i]) for i in range(2))
    batched = {"problems": len(maps), "seconds": round(t_batch, 3),
               "seconds_per_solve": round(t_batch / len(maps), 3),
               "plain_seconds_per_solve": round(t_plain, 3),
               "speedup_vs_plain": round(t_plain * len(maps) / t_batch, 2),
               "sample_bitwise_equal": bool(same),
               "note": "the first 8 of the sweep's 64 ratios, sector (0,0), one batch; plain: the "
                       "first 2 solved one after another"}
    log(f"[literal] batched: {len(maps)} solves in {t_batch:.2f} s ({t_batch / len(maps):.2f} s "
        f"each) against {t_plain:.2f} s alone: {batched['speedup_vs_plain']}x, bitwise {same}")
    t1 = solves["block1"]["seconds"]
    return {"workload": "two conserved components, 4x6 mesh each (48 orbitals), N_up = N_down = 4, "
                        "sector (0,0), inter-component ratio 1.3",
            "D": b.D, "rank_mode": b.rank_mode, "enumerate_s": round(t_enum, 3),
            "matvec_ms": round(ms_mv, 4), "solves": solves,
            "sweep_s_extrapolated": round(t1 * 24 * 64, 1),
            "batched": batched,
            "note": "single solves timed alone; the sweep figure is 1,536 such solves one after "
                    "another (extrapolated, not run)"}


def served_section(args):
    """The served stream through the continuous-batching server and through the plain
    one-at-a-time server, same requests, same HBM budget and queue bound."""
    from momentumed import serve as S
    reqs = S.served_stream(rate=50.0, duration=args.serve_seconds, seed=7)
    prefix = [r for r in reqs if r.arrival < args.serve_plain_seconds]
    out = {"stream": f"Poisson 50 req/s for {args.serve_seconds:g} s (seed 7), meshes "
                     f"{['%dx%d' % m for m in S.MESHES]}, fillings 1/3, 2/5 (non-integer N "
                     f"redrawn), one random sector, screening ~U(1,4), nev 4",
           "hbm_budget_bytes": 200 * 10 ** 9, "max_queue": 256, "drain_s": args.serve_drain,
           "plain_stream": f"the first {args.serve_plain_seconds:g} s of the same stream "
                           f"({len(prefix)} requests), same budget and queue bound"}
    out["policy"] = {"quantum_rows": 1 << 15, "big_share": args.serve_big_share,
                     "head_wait_s": args.serve_head_wait, "big_queue_limit": args.serve_big_queue}
    for mode, rq in (("batched", reqs), ("plain", prefix)):
        if mode == "plain" and args.serve_plain_seconds <= 0:
            continue
        srv = S.SolveServer(hbm_budget_bytes=200 * 10 ** 9, max_queue=256, mode=mode,
                            big_share=args.serve_big_share, head_wait_s=args.serve_head_wait,
                            big_queue_limit=args.serve_big_queue)
        try:
            rep = srv.run(rq, drain_s=args.serve_drain)
        finally:
            srv.close()
        summ = rep.summary()
        log(f"[served] {mode}: {summ['completed']}/{summ['offered']} done, "
            f"{summ['rejected']} rejected, {summ['unfinished']} unfinished, "
            f"{summ['solves_per_s']}/s, p50 {summ['p50_s']}, p99 {summ['p99_s']}, "
            f"p99 with unfinished {summ['p99_all_s']}")
        out[mode] = summ
    b = out["batched"]["solves_per_s"]
    p = out["plain"]["solves_per_s"] if "plain" in out else None
    out["speedup_vs_plain"] = round(b / p, 2) if b and p else None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="large", choices=["small", "medium", "large"])
    ap.add_argument("--no-solve", action="store_true",
                    help="skip timing the lowest-nev eigensolve of the sector")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default 512-row workgroups (skip the autotuner)")
    ap.add_argument("--no-block", action="store_true",
                    help="skip the block-matvec line (4 vectors per launch)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the sample point of the batched configuration as written")
    ap.add_argument("--batch-ratios", type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=120.0)
    ap.add_argument("--serve-plain-seconds", type=float, default=20.0,
                    help="the plain (one solve at a time) baseline serves this prefix")
    ap.add_argument("--serve-drain", type=float, default=30.0)
    ap.add_argument("--serve-big-share", type=float, default=0.5,
                    help="share of the stepped rows the oldest large running problem gets")
    ap.add_argument("--serve-head-wait", type=float, default=60.0,
                    help="seconds after which the queue's head is no longer passed")
    ap.add_argument("--serve-big-queue", type=int, default=128,
                    help="large requests (more rows than the step quantum) allowed to wait in "
                         "the queue (0: no limit beside the queue bound)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the small and medium configuration lines")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, run 3 plain device copies of one vector "
                         "(16*D bytes read + written each) to calibrate rocprof byte counters")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the package has no multi-GPU path (sectors and "
              "parameter points would shard one problem set per GPU). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import momentumed as me
    from momentumed import device as dev
    from momentumed import workloads as W
    from momentumed._lib import check

    w = W.get(args.workload)
    t0 = time.perf_counter()
    op = w.operator()
    t_terms = time.perf_counter() - t0
    log(f"[bench] {w.name}: {len(op.scats)} upper terms ({t_terms:.2f} s)")
    if not dev.activate_gpu_device("rocm"):
        raise RuntimeError("no GPU visible")
    ctx = dev.get_context(0)
    sysd = dev.system_struct(w.para.k_list, w.para.Gk, w.para.Nc_mix, w.para.Nc_conserve,
                             w.N_each)
    sector = me.preparation._momentum_list(w.para, w.N_each).index(w.sector)
    t0 = time.perf_counter()
    basis = dev.DeviceBasis.enumerate(ctx, sysd, sector)
    t_enum = time.perf_counter() - t0
    D = basis.D
    log(f"[bench] sector {w.sector}: D={D} enumerated in {t_enum:.3f} s, rank_mode={basis.rank_mode}")
    t0 = time.perf_counter()
    dop = dev.DeviceOperator.from_mboperator(ctx, op, w.para.bits)
    t_op = time.perf_counter() - t0
    nnz, nnz_off = ctypes.c_int64(), ctypes.c_int64()
    check(ctx.L.edm_connectivity_totals(ctx.h, dop.h, basis.h, ctypes.byref(nnz),
                                        ctypes.byref(nnz_off)), ctx.h)
    log(f"[bench] S_full={dop.S_full} nnz={nnz.value} nnz_off={nnz_off.value} "
        f"({nnz_off.value / D:.1f} off-diagonal per row)")

    x = torch.empty(D, dtype=torch.complex128, device="cuda:0")
    y = torch.empty_like(x)
    check(ctx.L.edm_fill_random(ctx.h, x.data_ptr(), D, 12345, 1), ctx.h)
    ctx.synchronize()
    x /= torch.linalg.vector_norm(x)
    torch.cuda.synchronize()
    lm = dev.ROCmLinearMap(dop, basis)
    if args.no_tune:
        tune_ms = None
    else:
        _, tune_ms = lm.tune()
    block = lm.kernel()
    log(f"[bench] matvec kernel {block:#x} ({_kernel_name(basis.rank_mode, block)}); tuner: "
        f"{tune_ms} ms per matvec")

    stream = torch.cuda.ExternalStream(ctx.stream, device="cuda:0")
    for _ in range(args.warmup):
        lm.matvec_device(y.data_ptr(), x.data_ptr())
    ctx.synchronize()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        lm.matvec_device(y.data_ptr(), x.data_ptr())
    ev1.record(stream)
    ev1.synchronize()
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_events = ev0.elapsed_time(ev1)
    ms_per = ms_events / args.steps
    log(f"[bench] {args.steps} matvecs: {ms_per:.3f} ms each (events), wall {wall:.3f} s")

    if args.copy_calibration:
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 copies of {16 * D} bytes")

    S = dop.S_full
    B_mv = D * 40 + nnz_off.value * 16 + S * 40
    achieved = B_mv / (ms_per * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": _kernel_name(basis.rank_mode, block),
                "bytes_model": "D*40 + nnz_off*16 + S_full*40", "algorithmic_bytes": B_mv}
    tr_path = os.path.join(ROOT, "profiles", f"matvec_{args.workload}_traffic.json")
    if os.path.exists(tr_path):
        tr = json.load(open(tr_path))
        roofline["traffic"] = tr.get("hbm_bytes_per_launch")
        roofline["traffic_source"] = os.path.relpath(tr_path, ROOT)

    out = {
        "metric": METRIC, "value": round(1e3 / ms_per, 3), "unit": "matvec/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": w.name, "mesh": [int(g) for g in w.para.Gk], "N": list(w.N_each),
                   "sector": list(w.sector), "D": D, "S_full": S, "nnz_offdiag": nnz_off.value,
                   "parallelism": "single-gpu", "matvec_workgroup": block & 0xFFFF,
                   "matvec_amplitudes_in_lds": bool(block & 0x10000),
                   "matvec_rank_low_table_in_lds": bool(block & 0x20000),
                   "tuner_ms": tune_ms},
        "roofline": roofline,
        "hbm_gbps_algorithmic": round(achieved, 1),
        "setup": {"terms_s": round(t_terms, 3), "enumerate_s": round(t_enum, 3),
                  "op_upload_s": round(t_op, 3), "rank_mode": basis.rank_mode},
        "wall_s_timed": round(wall, 4),
    }

    if not args.no_block:
        # the block kernel reads the low rank table from LDS where the single one does, for
        # bases of 2^18 rows and more (edm_matvec_block)
        out["block4"] = block_section(ctx, dop, basis, x, y, stream, args, nnz_off.value, S,
                                      ms_per, narrow=bool(block & 0x20000) and D >= (1 << 18))

    if not args.no_solve:
        log(f"[bench] eigensolve nev={w.nev} krylovdim={w.krylovdim} tol={w.tol}")
        ts = time.perf_counter()
        vals, _, info = dev.krylov_map_solve(lm, w.nev, krylovdim=w.krylovdim, tol=w.tol,
                                             want_vectors=False)
        secs = time.perf_counter() - ts
        mv_s = info["matvecs"] * ms_per * 1e-3
        out["solve"] = {"seconds": round(secs, 3), "nev": w.nev, "krylovdim": w.krylovdim,
                        "tol": w.tol, "eigenvalues": vals.tolist(), **info,
                        "matvec_seconds": round(mv_s, 3),
                        "matvec_share": round(mv_s / secs, 3)}
        log(f"[bench] solve: {out['solve']}")

    if not args.no_cpu_baseline:
        import numpy as np  # noqa: F811
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        amp, o, i = op.arrays()
        full = oracle.full_list(oracle.TermList(amp, o, i), True)
        basis_h = basis.download()
        x_h = x.cpu().numpy()
        y_h = y.cpu().numpy()
        nthreads, host = cpu_share()
        log(f"[bench] cpu baseline on {nthreads} threads (~{args.cpu_seconds:.0f} s): {host}")
        cb, err = cpu_baseline(w, basis_h, full, x_h, y_h, args.cpu_seconds, nthreads)
        cb["host"] = host
        out["cpu_baseline"] = cb
        out["parity_sampled_rows_max_rel_err"] = err
        log(f"[bench] cpu baseline {cb['value']:.3e} matvec/s; GPU/CPU rows max rel err {err:.2e}")
    else:
        out["cpu_baseline"] = None
    if not args.no_configs:
        out["configs"] = {n: config_section(n, args) for n in ("small", "medium")}
    if not args.no_batched:
        out["batched"] = batched_section(args)
        if not args.no_literal:
            out["batched"]["literal_sample"] = literal_section(args)
    if not args.no_served:
        out["served"] = served_section(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
